// mdgemm_kernels.hpp -- the reference's __global__ multi-double GEMM kernels,
// for HIP programs (compiled with hipcc) that launch them with <<<...>>> the
// way C++/TCcuda.cu:248-302 (matmulTCnt) does.  Defined in libmdgemm.so
// (csrc/ref_kernels.hip); link with -lmdgemm.
//
// Each kernel: A, B, C are n x n row-major matrices of interleaved p-doubles
// (p = 2, 4, 8, 16 for dd, qd, od, hd); the thread at
// (blockIdx.x*blockDim.x + threadIdx.x, blockIdx.y*blockDim.y + threadIdx.y)
// = (row, column) writes that entry of C = A*B, k ascending from zero, FMA
// TwoProd.  Any grid covering n x n is valid (threads outside return); the
// reference launches Gr(n/nfrag, n/nfrag) blocks of Bl(nfrag, nfrag).
// Results are bit-identical to mdg_vector_gemm(p, n, n, n, A, n, B, n, C, n,
// MDG_TWOPROD_FMA, stream).
#pragma once
#include <hip/hip_runtime.h>

// replaces C++/TCcuda.h:62 (double-double)
__global__ void ddmm(double *A, double *B, double *C, int n);
// replaces C++/TCcuda.h:63 (quad-double)
__global__ void qdmm(double *A, double *B, double *C, int n);
// replaces C++/TCcuda.h:64 (octo-double)
__global__ void odmm(double *A, double *B, double *C, int n);
// replaces C++/TCcuda.h:65 (hexa-double)
__global__ void hdmm(double *A, double *B, double *C, int n);
