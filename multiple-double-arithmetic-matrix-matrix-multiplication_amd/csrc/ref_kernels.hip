// ref_kernels.hip -- the reference's launchable GEMM kernels ddmm, qdmm, odmm,
// hdmm (C++/TCcuda.h:62-65, bodies C++/TCcuda.cu:119-246), exported as
// __global__ functions in the global namespace so a HIP program written
// against the reference launches them unchanged:
//
//     dim3 Gr(n / nfrag, n / nfrag), Bl(nfrag, nfrag);
//     qdmm<<<Gr, Bl>>>(A_d, B_d, C_d, n);
//
// Contract (the reference's): A, B, C are n x n row-major matrices of
// interleaved p-doubles (X[(r*n + c)*p + limb]); the thread at
// (blockIdx.x*blockDim.x + threadIdx.x, blockIdx.y*blockDim.y + threadIdx.y)
// writes C[r][c] = sum over k ascending of A[r][k]*B[k][c], starting from 0,
// with the multi-double mul and add of md_arith.hpp (FMA TwoProd, as the
// device code of this library everywhere).  Threads outside the n x n range
// return (the reference has no such check and needs nfrag to divide n).
//
// These kernels keep the reference's one-thread-per-output memory pattern
// so that any grid the reference's callers use is valid; mdg_vector_gemm
// (vector_gemm.hip) computes the same bits with LDS-staged tiles and is the
// fast form.
#include "../../include/mdgemm_kernels.hpp"
#include "md_arith.hpp"

namespace {

template <int L>
__device__ __forceinline__ void ref_entry(const double *__restrict__ A, const double *__restrict__ B,
                                          double *__restrict__ C, int n)
{
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long c = (long long)blockIdx.y * blockDim.y + threadIdx.y;
    if (r >= n || c >= n) return;
    double acc[L];
#pragma unroll
    for (int l = 0; l < L; ++l) acc[l] = 0.0;
    const double *a = A + r * n * L;
    const double *b = B + c * L;
    for (int k = 0; k < n; ++k) {
        double x[L], y[L], q[L];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            x[l] = a[(long long)k * L + l];
            y[l] = b[(long long)k * n * L + l];
        }
        mdg::md_mul<L, mdg::TP_FMA>(x, y, q);
        mdg::md_add<L>(acc, q, acc);
    }
    double *o = C + (r * n + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = acc[l];
}

}  // namespace

__global__ void ddmm(double *A, double *B, double *C, int n) { ref_entry<2>(A, B, C, n); }
__global__ void qdmm(double *A, double *B, double *C, int n) { ref_entry<4>(A, B, C, n); }
__global__ void odmm(double *A, double *B, double *C, int n) { ref_entry<8>(A, B, C, n); }
__global__ void hdmm(double *A, double *B, double *C, int n) { ref_entry<16>(A, B, C, n); }
