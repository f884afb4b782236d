// vector_gemm.hip -- the plain multi-double GEMM: one thread per output entry.
//
// Same arithmetic as the reference kernels ddmm/qdmm/odmm/hdmm
// (C++/TCcuda.cu:119-246): acc = 0; for k ascending acc = add(acc,
// mul(A[r][k], B[k][c])).  What differs is the memory path: the reference
// reads every limb straight from global memory with row-strided (uncoalesced)
// A loads; here a 256-thread workgroup stages a TM x TK tile of A and a
// TK x TN tile of B through LDS in limb-planar form ([limb][row][k]), loaded
// with coalesced reads of the interleaved layout, and every thread then runs
// its k-chain out of LDS.  The kernel is FP64-VALU bound (hundreds of
// dependent FP64 operations per multiply-add), so the tiles only have to
// keep the loads off the critical path.
//
// Each thread owns OUT outputs in the same row (columns tx, tx+TN/OUT, ...):
// their chains are independent, which gives the scheduler instruction-level
// parallelism to cover FP64 latency without raising the wave count.
#include "common.hpp"
#include "md_arith.hpp"

namespace mdg {

template <int L, int TP, int TM, int TN, int TK, int OUT>
__global__ __launch_bounds__(TM *(TN / OUT)) void vector_gemm_kernel(
    int m, int n, int k, const double *__restrict__ A, long long lda,
    const double *__restrict__ B, long long ldb, double *__restrict__ C, long long ldc)
{
    constexpr int NT = TM * (TN / OUT);   // threads per workgroup
    constexpr int TX = TN / OUT;          // threads along the columns
    __shared__ double As[L][TM][TK + 1];
    __shared__ double Bs[L][TK][TN + 1];

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const long long r0 = (long long)blockIdx.y * TM;
    const long long c0 = (long long)blockIdx.x * TN;

    double acc[OUT][L];
#pragma unroll
    for (int o = 0; o < OUT; ++o)
#pragma unroll
        for (int l = 0; l < L; ++l) acc[o][l] = 0.0;

    for (int k0 = 0; k0 < k; k0 += TK) {
        // A tile: TM rows, each TK entries of L limbs contiguous in memory
        for (int idx = tid; idx < TM * TK * L; idx += NT) {
            const int i = idx / (TK * L);
            const int rem = idx - i * (TK * L);
            const int kk = rem / L;
            const int l = rem - kk * L;
            const long long gr = r0 + i;
            const int gk = k0 + kk;
            double v = 0.0;
            if (gr < m && gk < k) v = A[(gr * lda + gk) * L + l];
            As[l][i][kk] = v;
        }
        // B tile: TK rows of k, each TN entries of L limbs contiguous
        for (int idx = tid; idx < TK * TN * L; idx += NT) {
            const int kk = idx / (TN * L);
            const int rem = idx - kk * (TN * L);
            const int j = rem / L;
            const int l = rem - j * L;
            const int gk = k0 + kk;
            const long long gc = c0 + j;
            double v = 0.0;
            if (gk < k && gc < n) v = B[((long long)gk * ldb + gc) * L + l];
            Bs[l][kk][j] = v;
        }
        __syncthreads();
        const int kmax = min(TK, k - k0);
        for (int kk = 0; kk < kmax; ++kk) {
            double a[L];
#pragma unroll
            for (int l = 0; l < L; ++l) a[l] = As[l][ty][kk];
#pragma unroll
            for (int o = 0; o < OUT; ++o) {
                double b[L], p[L];
#pragma unroll
                for (int l = 0; l < L; ++l) b[l] = Bs[l][kk][tx + o * TX];
                md_mul<L, TP>(a, b, p);
                md_add<L>(acc[o], p, acc[o]);
            }
        }
        __syncthreads();
    }
    const long long r = r0 + ty;
    if (r < m) {
#pragma unroll
        for (int o = 0; o < OUT; ++o) {
            const long long c = c0 + tx + o * TX;
            if (c < n) {
#pragma unroll
                for (int l = 0; l < L; ++l) C[(r * ldc + c) * L + l] = acc[o][l];
            }
        }
    }
}

// tile shape per limb count: (TM, TN, TK, OUT)
template <int L, int TP>
static int launch_vector(int m, int n, int k, const double *A, long long lda, const double *B,
                         long long ldb, double *C, long long ldc, hipStream_t s)
{
    constexpr int TM = 16, TN = (L <= 4) ? 32 : 16, TK = 16, OUT = (L <= 4) ? 2 : 1;
    dim3 grid((unsigned)((n + TN - 1) / TN), (unsigned)((m + TM - 1) / TM));
    dim3 block(TM * (TN / OUT));
    MDG_REQUIRE(grid.y <= 65535u, MDG_ERROR_NOT_SUPPORTED, "m=%d exceeds the grid limit", m);
    hipLaunchKernelGGL((vector_gemm_kernel<L, TP, TM, TN, TK, OUT>), grid, block, 0, s, m, n, k,
                       A, lda, B, ldb, C, ldc);
    MDG_LAUNCH_CHECK("vector_gemm_kernel");
    return MDG_SUCCESS;
}

int vector_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B,
                long long ldb, double *C, long long ldc, int twoprod, hipStream_t s)
{
    const bool dek = (twoprod == MDG_TWOPROD_DEKKER);
    switch (L) {
    case 2:
        return dek ? launch_vector<2, TP_DEKKER>(m, n, k, A, lda, B, ldb, C, ldc, s)
                   : launch_vector<2, TP_FMA>(m, n, k, A, lda, B, ldb, C, ldc, s);
    case 4:
        return dek ? launch_vector<4, TP_DEKKER>(m, n, k, A, lda, B, ldb, C, ldc, s)
                   : launch_vector<4, TP_FMA>(m, n, k, A, lda, B, ldb, C, ldc, s);
    case 8:
        return dek ? launch_vector<8, TP_DEKKER>(m, n, k, A, lda, B, ldb, C, ldc, s)
                   : launch_vector<8, TP_FMA>(m, n, k, A, lda, B, ldb, C, ldc, s);
    case 16:
        return dek ? launch_vector<16, TP_DEKKER>(m, n, k, A, lda, B, ldb, C, ldc, s)
                   : launch_vector<16, TP_FMA>(m, n, k, A, lda, B, ldb, C, ldc, s);
    }
    set_error("unsupported limb count L=%d (expected 2, 4, 8 or 16)", L);
    return MDG_ERROR_INVALID_VALUE;
}

}  // namespace mdg
