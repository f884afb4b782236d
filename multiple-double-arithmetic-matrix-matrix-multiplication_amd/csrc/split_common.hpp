// split_common.hpp -- pieces shared by the split-product GEMMs: the slice
// GEMMs (split_gemm.hip) and the residue GEMM (residue_gemm.hip).
#pragma once
#include <rocblas/rocblas.h>

#include <mutex>

#include "common.hpp"

namespace mdg {

constexpr int ZERO_SCALE = -4000;   // scale exponent of an all-zero row/column

// one rocBLAS handle per device, created (and its kernels loaded) on first use
struct BlasHandle {
    rocblas_handle h = nullptr;
    std::mutex mu;
};
BlasHandle *blas_handle();

// INT8 x INT8 -> INT32 strided-batched GEMM on this library's MFMA kernel
// (plain_i8_kernel), or on rocBLAS / hipBLASLt when mdg_set_int8_backend
// picks them (or the shape does not fit the kernel).  Column-major view:
// D_b (n x m, ldd) = Bt_b^T * A_b with Bt_b[c * ldb + kk], A_b[r * lda + kk],
// kk < depth (row-major: the m x n product of A_b and Bt_b^T).
// *engine_used: 2 the MFMA kernel, 1 rocBLAS, 3 hipBLASLt.
int i8_gemm_tn(int batch, int m, int n, int depth, const int8_t *Bt, long long ldb, long long sb, const int8_t *A,
               long long lda, long long sa, int *D, long long ldd, long long sd, hipStream_t s, int *engine_used);

// Scale exponents of A's rows (*E, m ints) and B's columns (*F, n ints) in
// scratch slots 0 and 1: every |entry| < 2^E, all-zero rows/columns get
// ZERO_SCALE.  With fixed_spread < 0 the exponent spreads are measured and
// read back (and non-finite limbs rejected); otherwise both spreads are
// fixed_spread and the stream stays asynchronous.
int scan_scales(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                int fixed_spread, hipStream_t s, int **E, int **F, int *spread_a, int *spread_b);

// D (m x n INT64, row stride n) = A (m x depth, row stride lda) * Bt^T
// (Bt: n x depth, row stride ldb) on the FP64 MFMA kernel, INT32 operands
// with |a b| * flushk <= 2^53 so every FP64 partial sum is exact.
int f64_gemm_i32(const int32_t *A, long long lda, const int32_t *Bt, long long ldb, int m, int n, int depth, int flushk,
                 long long *D, hipStream_t s);

inline long long round_up(long long x, long long q) { return (x + q - 1) / q * q; }

inline unsigned grid_for(long long nel, int bs) { return (unsigned)((nel + bs - 1) / bs); }

// Element stride between consecutive m x n result planes.  A power-of-two
// plane size puts the same offset of every plane on the same HBM channels,
// so a kernel that streams many planes at once (the reconstructions) runs
// at a fraction of the bandwidth; an odd multiple of 256 bytes of padding
// spreads the planes over the channels.
inline long long plane_stride(long long mn, int elem_bytes)
{
    return (round_up(mn * elem_bytes, 256) + 37 * 256) / elem_bytes;
}

}  // namespace mdg
