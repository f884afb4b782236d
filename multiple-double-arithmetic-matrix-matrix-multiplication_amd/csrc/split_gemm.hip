// split_gemm.hip -- split-product GEMM on the MFMA cores.
//
// C = A * B for m x k times k x n matrices of L-limb numbers, as
//   1. scan:        per row of A / column of B the scale exponent E (every
//                   |entry| < 2^E) and an exponent-spread estimate;
//   2. slice:       every entry cut into S balanced base-2^w digits of the
//                   fixed-point integer floor(x * 2^(W-2-E))  (slices.hpp);
//                   A slices land in A_all[r][t*kp + kk], B slices in
//                   Bt_all[c][(S-1-u)*kp + kk]: both operands K-contiguous,
//                   B's slices in reverse order;
//   3. diagonal GEMMs: for every anti-diagonal g < S,
//                   D_g = sum_{t+u=g} A_t * B_u
//                       = A_all[:, 0:(g+1)kp] * Bt_all[:, (S-1-g)kp : S kp]^T,
//                   one plain TN GEMM of depth (g+1)*kp per diagonal; all
//                   diagonals in one launch (blockIdx.z), deepest first;
//   4. reconstruct: C[r][c] = sum_g D_g[r][c] 2^(E_r+F_c+4-w(g+2)), formed
//                   exactly as an integer and folded into L limbs once with
//                   the TwoSum renormalisation the reference's squeeze uses
//                   (C++/TCcuda.cu:876-906; recon.hpp).
// This replaces the reference's split -> block-Toeplitz WMMA GEMM -> squeeze
// chain (C++/floatybits.cu:159-342, C++/TCcuda.cu:79-95, C++/TCtest.cu:21-46,
// C++/TCcuda.cu:876-958): the slice pairs with t+u <= S-1 are the non-zero
// blocks of the reference's Toeplitz operand, never materialised here.
//
// Three GEMM back ends share the slicing:
//   INT8  w = 8   v_mfma_i32_32x32x32_i8   exact INT32 sums while the depth K
//                 (the band kernel,        of one plane has K * 2^14 < 2^31;
//                 i8band.hpp, by default;  deeper bands are cut into K-chunks,
//                 the 256 x 256 diagonal   each writing one INT64 band plane,
//                 kernel or rocBLAS /      added in the reconstruction
//                 hipBLASLt on request)
//   BF16  w = 8   v_mfma_f32_32x32x16_bf16 exact FP32 sums (< 2^24), moved
//                 (band kernel; 16x16x32   into INT32 within each K-chunk
//                 in the diagonal kernel)
//   FP64  w <= 25 v_mfma_f64_16x16x4_f64   exact FP64 sums (< 2^53), widened
//                                          to INT64 every 2^(55-2w) k
#include <climits>
#include <cmath>

#include <rocblas/rocblas.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "blkslice.hpp"
#include "i8band.hpp"
#include "md_arith.hpp"
#include "recon.hpp"
#include "slices.hpp"
#include "split_common.hpp"

namespace mdg {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef int i16 __attribute__((ext_vector_type(16)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ scan --
// stats[0] (int): max spread over rows/cols, stats[1] (int): non-finite flag
__device__ __forceinline__ int scale_exponent(double s)
{
    // smallest E with s * (1 + 2^-40) < 2^E, s = sum of |limbs| of the entry
    if (s == 0.0) return INT_MIN;
    int ex;
    frexp(s * (1.0 + 0x1p-40), &ex);
    return ex;
}

// A rows: one wavefront per row
template <int L>
__global__ __launch_bounds__(256) void scan_rows_kernel(int m, int k, const double *__restrict__ A,
                                                        long long lda, int *__restrict__ E,
                                                        int *__restrict__ stats)
{
    const int lane = threadIdx.x & 63;
    const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= m) return;
    int emax = INT_MIN;
    double sum0 = 0.0;
    int bad = 0;
    for (int kk = lane; kk < k; kk += 64) {
        const double *x = A + (r * lda + kk) * L;
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const double v = x[l];
            bad |= !isfinite(v);
            s += fabs(v);
        }
        emax = max(emax, scale_exponent(s));
        sum0 += fabs(x[0]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        emax = max(emax, __shfl_xor(emax, o));
        sum0 += __shfl_xor(sum0, o);
        bad |= __shfl_xor(bad, o);
    }
    if (lane == 0) {
        E[r] = (emax == INT_MIN) ? ZERO_SCALE : emax;
        if (emax != INT_MIN && sum0 > 0.0) {
            atomicMax(&stats[0], exponent_spread(emax, sum0, k));
        }
        if (bad) atomicOr(&stats[1], 1);
    }
}

// B columns: 64 columns x 4 row groups per block, rows chunked over grid.y;
// partial results meet in E (atomicMax, pre-filled with INT_MIN); each
// block writes its column sums to part[blockIdx.y][c], which the finishing
// kernel adds in a fixed order (the same bits on every run).
template <int L>
__global__ __launch_bounds__(256) void scan_cols_kernel(int k, int n, const double *__restrict__ B,
                                                        long long ldb, int rows_per_block,
                                                        int *__restrict__ E, double *__restrict__ part,
                                                        int *__restrict__ stats)
{
    __shared__ int se[4][64];
    __shared__ double ss[4][64];
    const int cl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const long long c = (long long)blockIdx.x * 64 + cl;
    const int k0 = blockIdx.y * rows_per_block;
    const int k1 = min(k, k0 + rows_per_block);
    int emax = INT_MIN;
    double sum0 = 0.0;
    int bad = 0;
    if (c < n) {
        for (int kk = k0 + grp; kk < k1; kk += 4) {
            const double *x = B + ((long long)kk * ldb + c) * L;
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const double v = x[l];
                bad |= !isfinite(v);
                s += fabs(v);
            }
            emax = max(emax, scale_exponent(s));
            sum0 += fabs(x[0]);
        }
    }
    se[grp][cl] = emax;
    ss[grp][cl] = sum0;
    if (bad) atomicOr(&stats[1], 1);
    __syncthreads();
    if (grp == 0 && c < n) {
        int e = max(max(se[0][cl], se[1][cl]), max(se[2][cl], se[3][cl]));
        double s = (ss[0][cl] + ss[1][cl]) + (ss[2][cl] + ss[3][cl]);
        if (e != INT_MIN) atomicMax(&E[c], e);
        part[(long long)blockIdx.y * n + c] = s;
    }
}

// finish the column scan: zero columns get ZERO_SCALE, spread into stats[2]
// from the column sums of the nparts row chunks, added in chunk order
__global__ void finish_cols_kernel(int k, int n, int nparts, int *__restrict__ E,
                                   const double *__restrict__ part, int *__restrict__ stats)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int e = E[c];
    if (e == INT_MIN) {
        E[c] = ZERO_SCALE;
        return;
    }
    double s = 0.0;
    for (int y = 0; y < nparts; ++y) s += part[(long long)y * n + c];
    if (s > 0.0) {
        atomicMax(&stats[2], exponent_spread(e, s, k));
    }
}

__global__ void fill_int_kernel(int *p, long long nel, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nel) p[i] = v;
}

// ----------------------------------------------------------------- slice --
template <typename T>
__device__ __forceinline__ T digit_as(int d);
template <>
__device__ __forceinline__ int8_t digit_as<int8_t>(int d) { return (int8_t)d; }
template <>
__device__ __forceinline__ __bf16 digit_as<__bf16>(int d) { return (__bf16)(float)d; }
template <>
__device__ __forceinline__ int32_t digit_as<int32_t>(int d) { return d; }

// A: thread per entry (r, kk) of the padded mp x kp domain, kk fastest.
template <int L, typename T>
__global__ __launch_bounds__(256) void slice_a_kernel(int m, int k, int mp, int kp, const double *__restrict__ A,
                                                      long long lda, const int *__restrict__ E, int S, int w,
                                                      T *__restrict__ out)
{
    constexpr int NW = FixedWords<L>::NW;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)mp * kp) return;
    const long long r = idx / kp;
    const int kk = (int)(idx - r * kp);
    T *o = out + r * ((long long)S * kp) + kk;
    if (r >= m || kk >= k || E[r] == ZERO_SCALE) {
        for (int t = 0; t < S; ++t) o[(long long)t * kp] = digit_as<T>(0);
        return;
    }
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = A[(r * lda + kk) * L + l];
    uint32_t N[NW];
    to_fixed<L, NW>(x, S * w - 2 - E[r], N);
    fixed_digits<NW>(N, S, w, [&](int t, int d) { o[(long long)t * kp] = digit_as<T>(d); });
}

// B: thread per entry (kk, c) of the padded kp x np domain, kk fastest so
// the transposed, slice-reversed writes Bt[c][(S-1-u)*kp + kk] coalesce.
template <int L, typename T>
__global__ __launch_bounds__(256) void slice_b_kernel(int k, int n, int kp, int np, const double *__restrict__ B,
                                                      long long ldb, const int *__restrict__ F, int S, int w,
                                                      T *__restrict__ out)
{
    constexpr int NW = FixedWords<L>::NW;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)np * kp) return;
    const long long c = idx / kp;
    const int kk = (int)(idx - c * kp);
    T *o = out + c * ((long long)S * kp) + kk;
    if (c >= n || kk >= k || F[c] == ZERO_SCALE) {
        for (int u = 0; u < S; ++u) o[(long long)u * kp] = digit_as<T>(0);
        return;
    }
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = B[((long long)kk * ldb + c) * L + l];
    uint32_t N[NW];
    to_fixed<L, NW>(x, S * w - 2 - F[c], N);
    fixed_digits<NW>(N, S, w, [&](int u, int d) { o[(long long)(S - 1 - u) * kp] = digit_as<T>(d); });
}

// INT8 (or BF16) slices in the blocked layout of the band kernel
// (i8band.hpp), one tile of blk_rows_per_tile(S) rows x one k-block per
// workgroup (blkslice.hpp).  Grid: x = row tiles (rp / RT), y = k-blocks
// (kp / EB, EB = 128 / sizeof(T)).  TRANS: the operand is B (k x n, row
// stride ldx) and "row" r is column r of B.
template <int L, bool TRANS, typename T, int NW>
__global__ __launch_bounds__(256) void slice_blk_kernel(int rows, int k, int kp, const double *__restrict__ X,
                                                        long long ldx, const int *__restrict__ E, int S, int w,
                                                        int8_t *__restrict__ out)
{
    extern __shared__ __attribute__((aligned(16))) int8_t stage[];   // blk_stage_bytes(S, sizeof(T))
    slice_tile<L, TRANS, T, NW>(rows, k, kp, X, ldx, E, ZERO_SCALE, S, w, blockIdx.y,
                                blockIdx.x * blk_rows_per_tile(S, (int)sizeof(T)), out, stage);
}

// slices of an rp (padded) x kp operand in the blocked layout
template <int L, bool TRANS, typename T = int8_t>
static void launch_slice_blk(int rows, int k, int rp, int kp, const double *X, long long ldx, const int *E, int S,
                             int w, int8_t *out, hipStream_t s)
{
    const dim3 grid((unsigned)(rp / blk_rows_per_tile(S, (int)sizeof(T))), (unsigned)(kp / (BND_BK / (int)sizeof(T))));
    const int lds = blk_stage_bytes(S, (int)sizeof(T));
    with_slice_words<L>(S, w, [&](auto nw) {
        hipLaunchKernelGGL((slice_blk_kernel<L, TRANS, T, decltype(nw)::value>), grid, dim3(256), lds, s, rows, k, kp,
                           X, ldx, E, S, w, out);
    });
}

// --------------------------------------------------- INT8 diagonal GEMM --
// INT32 sums of one diagonal are exact while the chunk depth K satisfies
// K * 2^14 < 2^31, so a diagonal deeper than that is cut into K-chunks, each
// written to its own INT32 plane; the reconstruction adds the planes of one
// diagonal in INT64.  Plane p, in ascending diagonal order, is chunk c of
// diagonal g with nchunks(g) = ceil((g+1) kp / kmax).
__host__ __device__ inline int i8_chunks(int g, int kp, int kmax) { return ((g + 1) * kp + kmax - 1) / kmax; }

// plane index of diagonal g's first chunk: sum of the chunk counts before it
// (closed form when kmax = c * kp: sum_{j=1..g} ceil(j / c))
__host__ __device__ inline int i8_first_plane(int g, int kp, int kmax)
{
    if (kmax % kp == 0) {
        const int c = kmax / kp, q = g / c, r = g - q * c;
        return c * q * (q + 1) / 2 + r * (q + 1);
    }
    int first = 0;
    for (int h = 0; h < g; ++h) first += i8_chunks(h, kp, kmax);
    return first;
}

// deepest diagonal first: z-th plane of that order -> (g, c, first plane of g)
__device__ inline void i8_plane_of(int z, int S, int kp, int kmax, int &g, int &c)
{
    int rem = z;
    for (g = S - 1; g > 0; --g) {
        const int nc = i8_chunks(g, kp, kmax);
        if (rem < nc) break;
        rem -= nc;
    }
    c = rem;
}

// Hand-written backend: workgroup tile 256 x 256, 4 waves in 2 x 2, each
// 128 x 128 = 4 x 4 tiles of v_mfma_i32_32x32x32_i8 (256 INT32 sums per lane,
// all in the accumulation registers: one wave per SIMD).  Per k-block of 128
// bytes a wave reads 256 fragment rows from LDS for 64 MFMAs, so LDS traffic
// (reads and the block's writes) stays below the MFMA time.  k-blocks are
// double-buffered in LDS (2 x 2 x 256 rows x 144 bytes = 144 KiB; rows padded
// by 16 bytes: conflict-free 16-byte fragment reads): one barrier per
// k-block.  Operand buffers are padded (mp % 256 == np % 256 == kp % 128 ==
// 0), so the K loop carries no bounds checks.
__global__ __launch_bounds__(256) void diag_i8_kernel(const int8_t *__restrict__ Aall,
                                                      const int8_t *__restrict__ Ball, int S, int kp, int kmax,
                                                      int m, int n, int tiles_m, int tiles_n,
                                                      int *__restrict__ D)
{
    constexpr int BM = 256, BK = 128, LDR = BK + 16, TILE = BM * LDR, FP = BM / 32;
    __shared__ __attribute__((aligned(16))) int8_t lds[4 * TILE];   // A0 B0 A1 B1

    int g, c;
    i8_plane_of((int)blockIdx.z, S, kp, kmax, g, c);
    const int first = i8_first_plane(g, kp, kmax);
    const int k0 = c * kmax;
    const int k1 = min((g + 1) * kp, k0 + kmax);
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const long long ld = (long long)S * kp;
    const int8_t *Ab = Aall + (long long)tm * BM * ld + k0;
    const int8_t *Bb = Ball + (long long)tn * BM * ld + (long long)(S - 1 - g) * kp + k0;
    const int nkb = (k1 - k0) / BK;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 31, lk = (lane >> 5) * 16;

    i16 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

    // thread tid moves 16 bytes of rows tid/8 + 32 f, f < 8, of each operand:
    // a per-thread offset from uniform bases
    const int frow = tid >> 3, fseg = (tid & 7) * 16;
    const long long foff = frow * ld + fseg;
    const long long fstep = 32 * ld;
    const int soff = frow * LDR + fseg, sstep = 32 * LDR;
    // two register sets: the global loads run two k-blocks ahead of the
    // MFMAs (set kb & 1 receives block kb + 2 while set (kb + 1) & 1, holding
    // block kb + 1, waits to be written to LDS at the end of block kb)
    i4 ra[2][FP], rb[2][FP];
#define MDG_WIDE_FETCH(SET, KB)                                                     \
    do {                                                                            \
        const long long kofs_ = (long long)(KB) * BK;                               \
        _Pragma("unroll") for (int f = 0; f < FP; ++f)                              \
        {                                                                           \
            ra[SET][f] = *(const i4 *)(Ab + (f * fstep + kofs_) + foff);            \
            rb[SET][f] = *(const i4 *)(Bb + (f * fstep + kofs_) + foff);            \
        }                                                                           \
    } while (0)
    MDG_WIDE_FETCH(0, 0);
#pragma unroll
    for (int f = 0; f < FP; ++f) {
        *(i4 *)&lds[soff + f * sstep] = ra[0][f];
        *(i4 *)&lds[TILE + soff + f * sstep] = rb[0][f];
    }
    MDG_WIDE_FETCH(1, min(1, nkb - 1));
    __syncthreads();
    // fragments: lane l holds row l % 32, bytes 16 (l / 32) .. + 16 of a
    // 32-byte k-step, for A and B alike
    const int fra = (wm * 128 + lrow) * LDR + lk, frb = TILE + (wn * 128 + lrow) * LDR + lk;
    i4 fa[2][4], fb[2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) fa[0][i] = *(const i4 *)&lds[fra + i * 32 * LDR];
#pragma unroll
    for (int j = 0; j < 4; ++j) fb[0][j] = *(const i4 *)&lds[frb + j * 32 * LDR];
    // one k-block: SET receives block kb + 2 (the last one again near the
    // end: no branch, and nothing reads it); set SET ^ 1 goes to LDS
#define MDG_WIDE_BLOCK(SET)                                                                              \
    do {                                                                                                 \
        const int8_t *buf = lds + (kb & 1) * 2 * TILE;                                                   \
        int8_t *nbuf = lds + ((kb + 1) & 1) * 2 * TILE;                                                  \
        MDG_WIDE_FETCH(SET, min(kb + 2, nkb - 1));                                                       \
        __builtin_amdgcn_sched_barrier(0);                                                               \
        _Pragma("unroll") for (int t = 0; t < BK / 32 - 1; ++t)                                          \
        {                                                                                                \
            const int cur = t & 1, nxt = cur ^ 1;                                                        \
            _Pragma("unroll") for (int i = 0; i < 4; ++i) fa[nxt][i] =                                   \
                *(const i4 *)&buf[fra + i * 32 * LDR + (t + 1) * 32];                                    \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) fb[nxt][j] =                                   \
                *(const i4 *)&buf[frb + j * 32 * LDR + (t + 1) * 32];                                    \
            __builtin_amdgcn_sched_barrier(0);                                                           \
            _Pragma("unroll") for (int i = 0; i < 4; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)  \
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[cur][i], fb[cur][j], acc[i][j], 0, 0, 0); \
            __builtin_amdgcn_sched_barrier(0);                                                           \
        }                                                                                                \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)      \
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[1][i], fb[1][j], acc[i][j], 0, 0, 0);   \
        _Pragma("unroll") for (int f = 0; f < FP; ++f)                                                   \
        {                                                                                                \
            *(i4 *)&nbuf[soff + f * sstep] = ra[(SET) ^ 1][f];                                           \
            *(i4 *)&nbuf[TILE + soff + f * sstep] = rb[(SET) ^ 1][f];                                    \
        }                                                                                                \
        _Pragma("unroll") for (int q = 0; q < 16; ++q)                                                   \
        {                                                                                                \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                           \
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);                                           \
        }                                                                                                \
        __builtin_amdgcn_sched_barrier(0);                                                               \
        __syncthreads();                                                                                 \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) fa[0][i] = *(const i4 *)&nbuf[fra + i * 32 * LDR]; \
        _Pragma("unroll") for (int j = 0; j < 4; ++j) fb[0][j] = *(const i4 *)&nbuf[frb + j * 32 * LDR]; \
    } while (0)
    // k-steps 0-2 of a block read the fragments of step t + 1 from LDS while
    // the 16 MFMAs of step t run; step 3's MFMAs are interleaved with the 16
    // LDS writes of the next block into the spare buffer (last read before
    // the previous barrier); step 0 of the next block is read after the
    // barrier while the last MFMAs drain
    int kb = 0;
    for (; kb + 1 < nkb; kb += 2) {
        MDG_WIDE_BLOCK(0);
        ++kb;
        MDG_WIDE_BLOCK(1);
        --kb;
    }
    if (kb < nkb) MDG_WIDE_BLOCK(0);
#undef MDG_WIDE_BLOCK
#undef MDG_WIDE_FETCH
    // acc element v of tile (i, j) is row 8*(v/4) + 4*(lane/32) + v%4, col lane%32
    int *plane = D + (long long)(first + c) * m * n;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long col = (long long)tn * BM + wn * 128 + j * 32 + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long row = (long long)tm * BM + wm * 128 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (row < m && col < n) plane[row * n + col] = acc[i][j][v];
            }
        }
}

// --------------------------------------------------- BF16 diagonal GEMM --
// 64 bf16 (128 bytes) per k-block row and the 16x16x32 bf16 MFMA.  FP32 sums
// stay exact below 2^24, so every 16 k-blocks (1024 products of at most
// 2^14) they are moved into INT64 registers; to leave room for those the
// workgroup tile is 128 x 64 (waves 2 x 2, each 64 x 32).
__global__ __launch_bounds__(256) void diag_bf16_kernel(const __bf16 *__restrict__ Aall,
                                                        const __bf16 *__restrict__ Ball, int S, int kp,
                                                        int m, int n, int tiles_m, int tiles_n,
                                                        long long *__restrict__ Dout)
{
    constexpr int BM = 128, BN = 64, BK = 64, LDR = BK + 8;   // elements
    __shared__ __attribute__((aligned(16))) __bf16 As[BM * LDR];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[BN * LDR];

    const int g = S - 1 - (int)blockIdx.z;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const long long ld = (long long)S * kp;
    const __bf16 *Ab = Aall + (long long)tm * BM * ld;
    const __bf16 *Bb = Ball + (long long)tn * BN * ld + (long long)(S - 1 - g) * kp;
    const int Kd = (g + 1) * kp;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 15, lk = (lane >> 4) * 8;

    f4 acc[4][2];
    long long wide[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) wide[i][j][v] = 0;
        }

    // fetched one k-block ahead: thread tid moves 8 bf16 of rows tid/8 + 32 i
    const int frow = tid >> 3, fseg = (tid & 7) * 8;
    const long long foff = (long long)frow * ld + fseg, fstep = 32 * ld;
    const int soff = frow * LDR + fseg, sstep = 32 * LDR;
    uint4 ra0, ra1, ra2, ra3, rb0, rb1;
#define MDG_FETCH_BF16(kb)                                    \
    do {                                                      \
        ra0 = *(const uint4 *)(Ab + foff + (kb));             \
        ra1 = *(const uint4 *)(Ab + foff + fstep + (kb));     \
        ra2 = *(const uint4 *)(Ab + foff + 2 * fstep + (kb)); \
        ra3 = *(const uint4 *)(Ab + foff + 3 * fstep + (kb)); \
        rb0 = *(const uint4 *)(Bb + foff + (kb));             \
        rb1 = *(const uint4 *)(Bb + foff + fstep + (kb));     \
    } while (0)
    MDG_FETCH_BF16(0);
    int blk = 0;
    for (int kb = 0; kb < Kd; kb += BK, ++blk) {
        *(uint4 *)&As[soff] = ra0;
        *(uint4 *)&As[soff + sstep] = ra1;
        *(uint4 *)&As[soff + 2 * sstep] = ra2;
        *(uint4 *)&As[soff + 3 * sstep] = ra3;
        *(uint4 *)&Bs[soff] = rb0;
        *(uint4 *)&Bs[soff + sstep] = rb1;
        __syncthreads();
        if (kb + BK < Kd) MDG_FETCH_BF16(kb + BK);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            bf8 a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *(const bf8 *)&As[(wm * 64 + i * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *(const bf8 *)&Bs[(wn * 32 + j * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (((blk + 1) & 15) == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) wide[i][j][v] += (long long)acc[i][j][v];
                    acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
                }
        }
    }
    const long long plane = (long long)g * m * n;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const long long col = (long long)tn * BN + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long row = (long long)tm * BM + wm * 64 + i * 16 + (lane >> 4) * 4 + v;
                if (row < m && col < n) Dout[plane + row * n + col] = wide[i][j][v] + (long long)acc[i][j][v];
            }
        }
}

// --------------------------------------------------- FP64 diagonal GEMM --
// Workgroup tile 128 x 128, 8 waves in 2 x 4, each 64 x 32 = 4 x 2 MFMA
// 16x16x4 tiles; k-block of 32 staged in LDS as doubles (converted from the
// INT32 slices on the way in, halving the global traffic), fetched one
// k-block ahead into registers.  (76% of the FP64 matrix peak at the quad-
// double n = 4096 shape, against 52% for a 4-wave 128 x 64 tile.)
//   OUT 0: store FP64 sums          (limb convolution of the reference path)
//   OUT 1: add FP64 sums into C     (the reference's accumulate GEMM)
//   OUT 2: store INT64, moving the FP64 sums into INT64 every `flushk` k
//          (split path; the FP64 partial sums stay exact below 2^53)
// TA is the stored slice type (int32_t or double).  All loads are
// bounds-checked against (rows, Kd) so unpadded operands work.
template <typename TA, int OUT>
__global__ __launch_bounds__(512) void diag_f64_kernel(const TA *__restrict__ Aall, const TA *__restrict__ Ball,
                                                       int S, int kp, long long lda, long long ldb, int m, int n,
                                                       int tiles_m, int tiles_n, int flushk,
                                                       void *__restrict__ Dout, long long ldd)
{
    constexpr int NT = 512, BM = 128, BN = 128, BK = 32, LDR = BK + 1;
    constexpr int PA = BM * BK / NT, PB = BN * BK / NT;   // 8 each
    __shared__ double As[BM * LDR];
    __shared__ double Bs[BN * LDR];

    const int g = S - 1 - (int)blockIdx.z;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const long long row0 = (long long)tm * BM, col0 = (long long)tn * BN;
    const TA *Ab = Aall + row0 * lda;
    const TA *Bb = Ball + col0 * ldb + (long long)(S - 1 - g) * kp;
    const int Kd = (g + 1) * kp;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const int lrow = lane & 15, lk = lane >> 4;

    d4 acc[4][2];
    long long wide[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int v = 0; v < 4; ++v) wide[i][j][v] = 0;
        }

    // element idx = tid + it*512 of a tile is (row idx/32, k idx%32)
    TA ra[PA], rb[PB];   // raw slices; converted to double on the way into LDS
#pragma unroll
    for (int it = 0; it < PA; ++it) {
        const int idx = tid + it * NT, row = idx >> 5, kk = idx & 31;
        ra[it] = (row0 + row < m && kk < Kd) ? Ab[(long long)row * lda + kk] : (TA)0;
    }
#pragma unroll
    for (int it = 0; it < PB; ++it) {
        const int idx = tid + it * NT, row = idx >> 5, kk = idx & 31;
        rb[it] = (col0 + row < n && kk < Kd) ? Bb[(long long)row * ldb + kk] : (TA)0;
    }
    int kacc = 0;
    for (int kb = 0; kb < Kd; kb += BK) {
#pragma unroll
        for (int it = 0; it < PA; ++it) {
            const int idx = tid + it * NT;
            As[(idx >> 5) * LDR + (idx & 31)] = (double)ra[it];
        }
#pragma unroll
        for (int it = 0; it < PB; ++it) {
            const int idx = tid + it * NT;
            Bs[(idx >> 5) * LDR + (idx & 31)] = (double)rb[it];
        }
        __syncthreads();
        if (kb + BK < Kd) {
            const int kn = kb + BK;
#pragma unroll
            for (int it = 0; it < PA; ++it) {
                const int idx = tid + it * NT, row = idx >> 5, kk = idx & 31;
                ra[it] = (row0 + row < m && kn + kk < Kd) ? Ab[(long long)row * lda + kn + kk] : (TA)0;
            }
#pragma unroll
            for (int it = 0; it < PB; ++it) {
                const int idx = tid + it * NT, row = idx >> 5, kk = idx & 31;
                rb[it] = (col0 + row < n && kn + kk < Kd) ? Bb[(long long)row * ldb + kn + kk] : (TA)0;
            }
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            double a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[(wm * 64 + i * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[(wn * 32 + j * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if constexpr (OUT == 2) {
            kacc += BK;
            if (kacc >= flushk) {
                kacc = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) wide[i][j][v] += (long long)acc[i][j][v];
                        acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
                    }
            }
        }
    }
    const long long plane = (long long)g * m * ldd;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const long long col = col0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // v_mfma_f64_16x16x4f64: element v of lane l is row l/16 + 4v, col l%16
                const long long row = row0 + wm * 64 + i * 16 + (lane >> 4) + 4 * v;
                if (row < m && col < n) {
                    const long long o = plane + row * ldd + col;
                    if constexpr (OUT == 0)
                        ((double *)Dout)[o] = acc[i][j][v];
                    else if constexpr (OUT == 1)
                        ((double *)Dout)[o] += acc[i][j][v];
                    else
                        ((long long *)Dout)[o] = wide[i][j][v] + (long long)acc[i][j][v];
                }
            }
        }
}

// ------------------------------------------------------------ reconstruct --
// C[r][c] = sum_g D_g[r][c] 2^(E_r + F_c + 4 - w(g+2)), g = 0 most
// significant, formed exactly and folded into L limbs once (recon.hpp).
template <int L, int NW, int GS, typename PT>
__global__ __launch_bounds__(256) void reconstruct_kernel(int m, int n, int S, int w, const PT *__restrict__ D,
                                                          const int *__restrict__ E, const int *__restrict__ F,
                                                          double *__restrict__ C, long long ldc)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n) return;
    const long long r = idx / n, c = idx - r * n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = E[r], fc = F[c];
    if (er != ZERO_SCALE && fc != ZERO_SCALE) {
        const long long mn = (long long)m * n;
        diag_fold<L, NW, GS>(S, w, er + fc + 4 - w * (S + 1),
                             [&](int g) { return (long long)D[(long long)g * mn + idx]; }, y);
    }
    double *o = C + (r * ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

// INT8 planes (w = 8): with K-chunks, the planes of one diagonal are summed
// in INT64 first
template <int L, int NW, bool CHUNKED>
__global__ __launch_bounds__(256) void reconstruct_i8_kernel(int m, int n, int S, int kp, int kmax, int w,
                                                             const int *__restrict__ D, const int *__restrict__ E,
                                                             const int *__restrict__ F, double *__restrict__ C,
                                                             long long ldc)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n) return;
    const long long r = idx / n, c = idx - r * n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = E[r], fc = F[c];
    if (er != ZERO_SCALE && fc != ZERO_SCALE) {
        const long long mn = (long long)m * n;
        if constexpr (!CHUNKED) {
            // one INT32 plane per diagonal: the fold takes the sums as INT32
            diag_fold<L, NW, 3>(S, w, er + fc + 4 - w * (S + 1), [&](int g) { return D[(long long)g * mn + idx]; },
                                y);
        } else {
            diag_fold<L, NW, 3>(S, w, er + fc + 4 - w * (S + 1),
                                [&](int g) {
                                    const int first = i8_first_plane(g, kp, kmax);
                                    const int nc = i8_chunks(g, kp, kmax);
                                    long long v = 0;
                                    for (int q = 0; q < nc; ++q) v += D[(long long)(first + q) * mn + idx];
                                    return v;
                                },
                                y);
        }
    }
    double *o = C + (r * ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

// Band planes of the band kernels (i8band.hpp, INT8 bands of G = 4 or BF16
// bands of 3; w = 8): one INT64 plane per (band, K-chunk), bp.nch[q] chunks
// of band q (bnd_plane_table), unchunked when every band is one chunk
template <int L, int NWB, int G, bool CHUNKED>
__global__ __launch_bounds__(256) void reconstruct_band_kernel(int m, int n, int S, BandPlanes bp,
                                                               const long long *__restrict__ D,
                                                               const int *__restrict__ E, const int *__restrict__ F,
                                                               double *__restrict__ C, long long ldc)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n) return;
    const long long r = idx / n, c = idx - r * n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = E[r], fc = F[c];
    if (er != ZERO_SCALE && fc != ZERO_SCALE) {
        const long long mn = (long long)m * n;
        band_fold<L, NWB, G, CHUNKED>(
            S, er + fc + 4 - 8 * (S + 1),
            [&](int plane) { return D[(long long)plane * mn + idx]; }, [&](int q) { return (int)bp.nch[q]; },
            bp.total, y);
    }
    double *o = C + (r * ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

template <int L, int G>
static int launch_band_recon(int m, int n, int S, int nkb, int cap, const long long *D, const int *E, const int *F,
                             double *C, long long ldc, hipStream_t s)
{
    const dim3 grid((unsigned)(((long long)m * n + 255) / 256));
    const BandPlanes bp = bnd_plane_table(S, G, nkb, cap);
    MDG_REQUIRE(bp.nb <= 96, MDG_ERROR_RANGE, "%d slices: more bands than the reconstruction's table holds", S);
    const bool chunked = bp.total > bp.nb;   // a band of several K-chunks
    return with_band_words<L, G>(S, [&](auto nwb) {
        constexpr int NWB = decltype(nwb)::value;
        if (chunked)
            hipLaunchKernelGGL((reconstruct_band_kernel<L, NWB, G, true>), grid, dim3(256), 0, s, m, n, S, bp, D, E,
                               F, C, ldc);
        else
            hipLaunchKernelGGL((reconstruct_band_kernel<L, NWB, G, false>), grid, dim3(256), 0, s, m, n, S, bp, D, E,
                               F, C, ldc);
    });
}

// ------------------------------------------------------- rocBLAS backend --
// The diagonal GEMMs are plain INT8 x INT8 -> INT32 GEMMs; rocBLAS runs each
// K-chunk plane as one gemm_ex call (column-major view: D_p^T (n x m) =
// Bt_chunk^T-op (n x K) * A_chunk (K x m), lda = ldb = S*kp).
static std::mutex g_blas_mu;
static std::map<int, BlasHandle *> g_blas;

BlasHandle *blas_handle()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_blas_mu);
    auto it = g_blas.find(dev);
    if (it != g_blas.end()) return it->second;
    rocblas_initialize();   // load the GEMM kernels now, not inside a timed call
    BlasHandle *b = new BlasHandle();
    if (rocblas_create_handle(&b->h) != rocblas_status_success) {
        delete b;
        return nullptr;
    }
    g_blas[dev] = b;
    return b;
}

// Slice-plane engines (mdg_set_int8_backend): 0 and 2 -- the band kernels
// (i8band.hpp) for INT8 and BF16 slices, or the single-diagonal kernels
// where a band kernel cannot run (more than 32767 k-blocks, or more bands
// than the reconstruction's chunk table holds); 4 -- the single-diagonal
// kernels (256 x 256 INT8, INT64-plane BF16); 1 rocBLAS and 3 hipBLASLt --
// the vendor libraries for INT8 slices, kept only as comparison engines.
enum { I8E_BAND = 2, I8E_LIB = 1, I8E_TILE256 = 4 };

// Band items of one problem (band_items: every item G wide, the leftover
// diagonals as a shifted band), heaviest first; false where the band
// kernels cannot run it (k-block indices are shorts, the reconstruction's
// chunk table holds 96 bands).
static bool build_band_items(int S, int G, int nkb, int cap, std::vector<BandItem> &items)
{
    items.clear();
    if (nkb > 32767 || (S + G - 1) / G > 96) return false;
    std::vector<BandWork> v;
    band_items(S, G, nkb, cap, v);
    std::stable_sort(v.begin(), v.end(), [](const BandWork &x, const BandWork &y) { return x.work > y.work; });
    for (const BandWork &w : v) items.push_back(w.it);
    return true;
}

// The band kernel over the problem's items, BND_MAX_ITEMS items per launch
// (the table is a kernel argument); every item writes its own plane, so the
// launches are independent.  BF16 selects the kernel over blocked BF16
// slices (bands of 3 diagonals, nkb = kp / 64), else INT8 (4, kp / 128).
template <bool BF16>
static int launch_band(const std::vector<BandItem> &items, const int8_t *Aall, const int8_t *Ball, int S, int nkb,
                       int cap, int tiles_m, int tiles_n, int m, int n, long long *D, hipStream_t s)
{
    for (size_t i0 = 0; i0 < items.size(); i0 += BND_MAX_ITEMS) {
        BandTable tab{};
        tab.count = (int)std::min(items.size() - i0, (size_t)BND_MAX_ITEMS);
        for (int i = 0; i < tab.count; ++i) tab.item[i] = items[i0 + i];
        const dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)tab.count);
        if (BF16) {
            hipLaunchKernelGGL((band_bf16_kernel<3, 2, 1>), grid, dim3(256), 0, s, Aall, Ball, S, nkb, cap, tiles_m,
                               tiles_n, m, n, tab, D);
            MDG_LAUNCH_CHECK("band_bf16_kernel");
        } else {
            hipLaunchKernelGGL((band_i8_kernel<4, 2, 1>), grid, dim3(256), 0, s, Aall, Ball, S, nkb, cap, tiles_m,
                               tiles_n, m, n, tab, D);
            MDG_LAUNCH_CHECK("band_i8_kernel");
        }
    }
    return MDG_SUCCESS;
}

// every K-chunk plane as one INT8 GEMM on the library engines (rocBLAS or
// hipBLASLt, autotuned per plane shape); *engine: the engine of the last plane
static int i8_planes_library(const int8_t *Aall, const int8_t *Ball, int S, int kp, int kmax, int m, int n, int *D,
                             hipStream_t s, int *engine)
{
    const long long ld = (long long)S * kp;
    const long long mn = (long long)m * n;
    int p = 0;
    for (int g = 0; g < S; ++g) {
        const int K = (g + 1) * kp;
        for (int k0 = 0; k0 < K; k0 += kmax, ++p) {
            const int kk = std::min(kmax, K - k0);
            const int rc = i8_gemm_tn(1, m, n, kk, Ball + (long long)(S - 1 - g) * kp + k0, ld, 0, Aall + k0, ld, 0,
                                      D + (long long)p * mn, n, mn, s, engine);
            if (rc) return rc;
        }
    }
    return MDG_SUCCESS;
}

// ---------------------------------------------------------------- driver --
template <typename K>
static inline unsigned blocks_for(long long nel, int bs) { return (unsigned)((nel + bs - 1) / bs); }

struct Plan {
    int method, L, m, n, k;
    int S, w, W;
    int kp, mp, np;
    bool wide;        // INT8: widened accumulation
    int flush_blocks; // FP64 / INT8
    int spread_a, spread_b;
    int backend;      // INT8: I8E_BAND, I8E_TILE256 or I8E_LIB
};

// Choose slice width and count for a precision target and the measured
// spreads (see slices.hpp for the error terms): the dominant error of the
// product is (S+1) * L * 2^(2 - W + spread_a + spread_b) relative to
// sum_k |a||b|, so W = target + spreads + log2((S+1) L) + 2.
static int plan_slices(Plan &p, int target_bits)
{
    const int base = target_bits + p.spread_a + p.spread_b + 2;
    auto need_W = [&](int S) {
        const double extra = std::ceil(std::log2((double)(S + 1) * p.L));
        return base + (int)extra;
    };
    if (p.method == MDG_METHOD_INT8_OZAKI || p.method == MDG_METHOD_BF16_OZAKI) {
        p.w = 8;
        int S = 4;
        while (S * 8 < need_W(S)) ++S;
        p.S = S;
    } else {
        // FP64: widest w with (2w-2) + log2(S * kp) <= 62 (INT64 diagonal
        // sums) and one 16-deep k-block exact in FP64 (16 * 2^(2w-2) <= 2^53)
        p.w = 0;
        for (int w = 25; w >= 8; --w) {   // 2^(55-2w) >= 32: one k-block exact
            int S = 2;
            while (S * w < need_W(S)) ++S;
            const double bits = (2.0 * w - 2.0) + std::log2((double)S * p.kp);
            if (bits <= 62.0) {
                p.w = w;
                p.S = S;
                break;
            }
        }
        if (p.w == 0) return MDG_ERROR_RANGE;
        // FP64 partial sums exact while count * 2^(2w-2) <= 2^53
        // (for FP64, flush_blocks holds the widening period in k)
        p.flush_blocks = (int)(1ll << (55 - 2 * p.w));
    }
    p.W = p.S * p.w;
    return MDG_SUCCESS;
}

// slice count of the INT8 path for a given precision target and spreads
// (the batched pipeline sizes its slices with this, without a read-back)
int plan_int8_slices(int L, int target_bits, int spread_a, int spread_b)
{
    Plan p{};
    p.method = MDG_METHOD_INT8_OZAKI;
    p.L = L;
    p.spread_a = spread_a;
    p.spread_b = spread_b;
    p.kp = 128;
    return plan_slices(p, target_bits) ? -1 : p.S;
}

// stats of the last scan on this thread (the guard of a fixed-spread call reads them)
static thread_local int *g_last_stats = nullptr;
// where the guard of the next fixed-spread call stores its flags (a batch's
// per-problem flag, set by the batched loop; NULL: not stored)
static thread_local int *g_guard_flag = nullptr;

void set_guard_flag(int *flag) { g_guard_flag = flag; }

// A call with a caller-fixed spread allowance (no read-back) checks on the
// device what the scans measured: bit 0 = an input limb is Inf or NaN, bit 1
// = the exponent spread of A's rows or B's columns exceeds the allowance, so
// the slices were too few for the precision target.  A flagged result is
// filled with NaN (no silent inaccuracy) and the flags go to *flag.
__global__ void guard_kernel(const int *__restrict__ stats, int allow, int m, int n, double *__restrict__ C,
                             long long ldc, int L, int *__restrict__ flag)
{
    const int f = (stats[1] ? 1 : 0) | ((stats[0] > allow || stats[2] > allow) ? 2 : 0);
    if (flag && blockIdx.x == 0 && threadIdx.x == 0) *flag = f;
    if (!f) return;
    const long long nel = (long long)m * n * L;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nel; i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / ((long long)n * L), rem = i - r * n * L;
        C[r * ldc * L + rem] = __builtin_nan("");
    }
}

template <int L>
static int scan_scales_l(int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                         int fixed_spread, hipStream_t s, int **Eo, int **Fo, int *spread_a, int *spread_b)
{
    // scratch: E (m), F (n), stats (8) in slot 0; the column sums of the
    // row chunks (nparts x n) in slot 1
    int *E = (int *)scratch(s, sizeof(int) * (size_t)(m + n + 8), 0);
    MDG_REQUIRE(E, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    int *F = E + m;
    int *stats = F + n;
    g_last_stats = stats;
    const int rpb = 256, nparts = (k + rpb - 1) / rpb;
    double *part = (double *)scratch(s, sizeof(double) * (size_t)nparts * n, 1);
    MDG_REQUIRE(part, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    hipLaunchKernelGGL(fill_int_kernel, dim3(blocks_for<int>(n, 256)), dim3(256), 0, s, F, (long long)n, INT_MIN);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, stats, 8ll, 0);
    hipLaunchKernelGGL((scan_rows_kernel<L>), dim3((m + 3) / 4), dim3(256), 0, s, m, k, A, lda, E, stats);
    hipLaunchKernelGGL((scan_cols_kernel<L>), dim3((n + 63) / 64, nparts), dim3(256), 0, s, k, n, B, ldb, rpb, F,
                       part, stats);
    hipLaunchKernelGGL(finish_cols_kernel, dim3(blocks_for<int>(n, 256)), dim3(256), 0, s, k, n, nparts, F, part,
                       stats);
    MDG_LAUNCH_CHECK("scan kernels");
    if (fixed_spread < 0) {
        int hstats[4];
        MDG_HIP_CHECK(hipMemcpyAsync(hstats, stats, sizeof(hstats), hipMemcpyDeviceToHost, s));
        MDG_HIP_CHECK(hipStreamSynchronize(s));
        MDG_REQUIRE(hstats[1] == 0, MDG_ERROR_NON_FINITE, "input contains Inf or NaN limbs");
        *spread_a = hstats[0];
        *spread_b = hstats[2];
    } else {
        // caller-supplied allowance: no read-back, the stream stays asynchronous
        *spread_a = fixed_spread;
        *spread_b = fixed_spread;
    }
    *Eo = E;
    *Fo = F;
    return MDG_SUCCESS;
}

int scan_scales(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                int fixed_spread, hipStream_t s, int **E, int **F, int *spread_a, int *spread_b)
{
    switch (L) {
    case 2: return scan_scales_l<2>(m, n, k, A, lda, B, ldb, fixed_spread, s, E, F, spread_a, spread_b);
    case 4: return scan_scales_l<4>(m, n, k, A, lda, B, ldb, fixed_spread, s, E, F, spread_a, spread_b);
    case 8: return scan_scales_l<8>(m, n, k, A, lda, B, ldb, fixed_spread, s, E, F, spread_a, spread_b);
    case 16: return scan_scales_l<16>(m, n, k, A, lda, B, ldb, fixed_spread, s, E, F, spread_a, spread_b);
    }
    set_error("unsupported limb count L=%d", L);
    return MDG_ERROR_INVALID_VALUE;
}

template <int L>
static int run_split_gemm(Plan &p, int target_bits, int fixed_spread, const double *A, long long lda,
                          const double *B, long long ldb, double *C, long long ldc, hipStream_t s,
                          mdg_gemm_info *info)
{
    const int m = p.m, n = p.n, k = p.k;
    StageTimer tm(s, profiling_on());
    tm.mark(0);
    int *E = nullptr, *F = nullptr;
    int rc = scan_scales_l<L>(m, n, k, A, lda, B, ldb, fixed_spread, s, &E, &F, &p.spread_a, &p.spread_b);
    if (rc) return rc;
    tm.mark(1);

    p.kp = (int)round_up(k, 128);
    p.mp = (int)round_up(m, 128);
    p.np = (int)round_up(n, 128);
    rc = plan_slices(p, target_bits);
    if (rc) {
        set_error("no slice width fits k=%d at %d target bits", k, target_bits);
        return rc;
    }
    MDG_REQUIRE(p.W <= 32 * FixedWords<L>::NW - 1, MDG_ERROR_RANGE,
                "needs %d fixed-point bits, more than the %d this limb count holds", p.W,
                32 * FixedWords<L>::NW - 1);
    const int S = p.S, w = p.w, kp = p.kp, mp = p.mp, np = p.np;
    const long long mn = (long long)m * n;
    const int tiles_m = mp / 128;
    const unsigned zdim = (unsigned)S;
    MDG_REQUIRE(zdim <= 65535u, MDG_ERROR_RANGE, "too many slices (%d)", S);

    int lib_engine_used = 1;
    // inside a no-allocation call only the hand-written kernels run
    const int forced = noalloc_active() ? 0 : int8_backend();
    std::vector<BandItem> items;
    if (p.method == MDG_METHOD_INT8_OZAKI) {
        // INT32 planes exact while chunk depth K * 2^14 < 2^31
        const int kmax = (kp <= 131071) ? (131071 / kp) * kp : 130944;
        int nplanes = 0;
        for (int g = 0; g < S; ++g) nplanes += i8_chunks(g, kp, kmax);
        p.wide = nplanes > S;
        p.flush_blocks = kmax;
        int engine = (forced == 1 || forced == 3) ? I8E_LIB : (forced == 4 ? I8E_TILE256 : I8E_BAND);
        // the band kernel cuts any depth into exact INT32 K-chunks
        if (engine == I8E_BAND && !build_band_items(S, 4, kp / 128, BND_CAP_I8, items)) engine = I8E_TILE256;
        // the band kernel writes one INT64 plane per (band, chunk), the
        // other engines one INT32 plane per (diagonal, chunk)
        const int nbp = (engine == I8E_BAND) ? bnd_band_planes(S, 4, kp / 128, BND_CAP_I8) : 0;
        void *D = scratch(s, (engine == I8E_BAND) ? (size_t)nbp * mn * 8 : (size_t)nplanes * mn * 4, 4);
        MDG_REQUIRE(D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
        if (engine == I8E_BAND) {
            // blocked slices, 128-row tiles
            const int mpb = (int)round_up(m, 128), npb = (int)round_up(n, 128);
            int8_t *Aall = (int8_t *)scratch(s, (size_t)mpb * S * kp, 2);
            int8_t *Ball = (int8_t *)scratch(s, (size_t)npb * S * kp, 3);
            MDG_REQUIRE(Aall && Ball, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
            launch_slice_blk<L, false>(m, k, mpb, kp, A, lda, E, S, w, Aall, s);
            launch_slice_blk<L, true>(n, k, npb, kp, B, ldb, F, S, w, Ball, s);
            tm.mark(2);
            const int tiles_mb = mpb / 128, tiles_nb = npb / 128;
            rc = launch_band<false>(items, Aall, Ball, S, kp / 128, BND_CAP_I8, tiles_mb, tiles_nb, m, n,
                                    (long long *)D, s);
            if (rc) return rc;
        } else {
            // row-major K-contiguous slices: A_all[r][t kp + kk], Bt_all[c][(S-1-u) kp + kk]
            const int mp8 = (int)round_up(m, 256), np8 = (int)round_up(n, 256);
            int8_t *Aall = (int8_t *)scratch(s, (size_t)mp8 * S * kp, 2);
            int8_t *Ball = (int8_t *)scratch(s, (size_t)np8 * S * kp, 3);
            MDG_REQUIRE(Aall && Ball, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
            hipLaunchKernelGGL((slice_a_kernel<L, int8_t>), dim3(blocks_for<int>((long long)mp8 * kp, 256)),
                               dim3(256), 0, s, m, k, mp8, kp, A, lda, E, S, w, Aall);
            hipLaunchKernelGGL((slice_b_kernel<L, int8_t>), dim3(blocks_for<int>((long long)np8 * kp, 256)),
                               dim3(256), 0, s, k, n, kp, np8, B, ldb, F, S, w, Ball);
            tm.mark(2);
            if (engine == I8E_LIB) {
                rc = i8_planes_library(Aall, Ball, S, kp, kmax, m, n, (int *)D, s, &lib_engine_used);
                if (rc) return rc;
            } else {
                MDG_REQUIRE(nplanes <= 65535, MDG_ERROR_RANGE, "too many INT8 planes (%d)", nplanes);
                const int tiles_m8 = mp8 / 256, tiles_n8 = np8 / 256;
                dim3 grid(tiles_m8 * tiles_n8, 1, (unsigned)nplanes);
                hipLaunchKernelGGL(diag_i8_kernel, grid, dim3(256), 0, s, Aall, Ball, S, kp, kmax, m, n, tiles_m8,
                                   tiles_n8, (int *)D);
                MDG_LAUNCH_CHECK("diag_i8_kernel");
            }
        }
        p.backend = engine;
        tm.mark(3);
        if (engine == I8E_BAND)
            rc = launch_band_recon<L, 4>(m, n, S, kp / 128, BND_CAP_I8, (const long long *)D, E, F, C, ldc, s);
        else
            rc = with_diag_words<L>(S, w, [&](auto nw) {
                constexpr int NW = decltype(nw)::value;
                if (nplanes > S)
                    hipLaunchKernelGGL((reconstruct_i8_kernel<L, NW, true>), dim3(blocks_for<int>(mn, 256)),
                                       dim3(256), 0, s, m, n, S, kp, kmax, w, (const int *)D, E, F, C, ldc);
                else
                    hipLaunchKernelGGL((reconstruct_i8_kernel<L, NW, false>), dim3(blocks_for<int>(mn, 256)),
                                       dim3(256), 0, s, m, n, S, kp, kmax, w, (const int *)D, E, F, C, ldc);
            });
        if (rc) return rc;
    } else if (p.method == MDG_METHOD_BF16_OZAKI && forced != 4 &&
               build_band_items(S, 3, kp / 64, BND_CAP_BF16, items)) {
        // band kernel over blocked BF16 slices: the same exact integer sums
        // as the INT8 path (FP32 partial sums moved into INT32 every 1024
        // products), as INT64 band planes of 3 diagonals (192 FP32 sums and
        // 192 INT32 totals per lane); beyond 32767 k-blocks the
        // single-diagonal kernel below runs
        const int kmax = (kp <= 131071) ? (131071 / kp) * kp : 131008;
        int nplanes = 0;
        for (int g = 0; g < S; ++g) nplanes += i8_chunks(g, kp, kmax);
        p.wide = nplanes > S;
        p.flush_blocks = kmax;
        long long *D = (long long *)scratch(s, (size_t)bnd_band_planes(S, 3, kp / 64, BND_CAP_BF16) * mn * 8, 4);
        const int mpb = (int)round_up(m, 128), npb = (int)round_up(n, 128);
        int8_t *Aall = (int8_t *)scratch(s, (size_t)mpb * S * kp * 2, 2);
        int8_t *Ball = (int8_t *)scratch(s, (size_t)npb * S * kp * 2, 3);
        MDG_REQUIRE(Aall && Ball && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
        launch_slice_blk<L, false, __bf16>(m, k, mpb, kp, A, lda, E, S, w, Aall, s);
        launch_slice_blk<L, true, __bf16>(n, k, npb, kp, B, ldb, F, S, w, Ball, s);
        tm.mark(2);
        const int tiles_mb = mpb / 128, tiles_nb = npb / 128;
        rc = launch_band<true>(items, Aall, Ball, S, kp / 64, BND_CAP_BF16, tiles_mb, tiles_nb, m, n, D, s);
        if (rc) return rc;
        p.backend = I8E_BAND;
        tm.mark(3);
        rc = launch_band_recon<L, 3>(m, n, S, kp / 64, BND_CAP_BF16, D, E, F, C, ldc, s);
        if (rc) return rc;
    } else if (p.method == MDG_METHOD_BF16_OZAKI) {
        // single-diagonal kernel: INT64 diagonal planes
        p.wide = true;
        p.backend = I8E_TILE256;
        p.flush_blocks = 16;
        __bf16 *Aall = (__bf16 *)scratch(s, (size_t)mp * S * kp * 2, 2);
        __bf16 *Ball = (__bf16 *)scratch(s, (size_t)np * S * kp * 2, 3);
        long long *D = (long long *)scratch(s, (size_t)S * mn * 8, 4);
        MDG_REQUIRE(Aall && Ball && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
        hipLaunchKernelGGL((slice_a_kernel<L, __bf16>), dim3(blocks_for<int>((long long)mp * kp, 256)), dim3(256), 0,
                           s, m, k, mp, kp, A, lda, E, S, w, Aall);
        hipLaunchKernelGGL((slice_b_kernel<L, __bf16>), dim3(blocks_for<int>((long long)np * kp, 256)), dim3(256), 0,
                           s, k, n, kp, np, B, ldb, F, S, w, Ball);
        tm.mark(2);
        const int tiles_n = np / 64;
        dim3 grid(tiles_m * tiles_n, 1, zdim);
        hipLaunchKernelGGL(diag_bf16_kernel, grid, dim3(256), 0, s, Aall, Ball, S, kp, m, n, tiles_m, tiles_n, D);
        MDG_LAUNCH_CHECK("diag_bf16_kernel");
        tm.mark(3);
        rc = with_diag_words<L>(S, w, [&](auto nw) {
            with_group(w, [&](auto gs) {
                hipLaunchKernelGGL((reconstruct_kernel<L, decltype(nw)::value, decltype(gs)::value, long long>),
                                   dim3(blocks_for<int>(mn, 256)), dim3(256), 0, s, m, n, S, w,
                                   (const long long *)D, E, F, C, ldc);
            });
        });
        if (rc) return rc;
    } else {
        p.wide = true;
        int32_t *Aall = (int32_t *)scratch(s, (size_t)mp * S * kp * 4, 2);
        int32_t *Ball = (int32_t *)scratch(s, (size_t)np * S * kp * 4, 3);
        long long *D = (long long *)scratch(s, (size_t)S * mn * 8, 4);
        MDG_REQUIRE(Aall && Ball && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
        hipLaunchKernelGGL((slice_a_kernel<L, int32_t>), dim3(blocks_for<int>((long long)mp * kp, 256)), dim3(256), 0,
                           s, m, k, mp, kp, A, lda, E, S, w, Aall);
        hipLaunchKernelGGL((slice_b_kernel<L, int32_t>), dim3(blocks_for<int>((long long)np * kp, 256)), dim3(256), 0,
                           s, k, n, kp, np, B, ldb, F, S, w, Ball);
        tm.mark(2);
        const int tiles_n = np / 128;
        dim3 grid(tiles_m * tiles_n, 1, zdim);
        hipLaunchKernelGGL((diag_f64_kernel<int32_t, 2>), grid, dim3(512), 0, s, Aall, Ball, S, kp,
                           (long long)S * kp, (long long)S * kp, m, n, tiles_m, tiles_n, p.flush_blocks, (void *)D,
                           (long long)n);
        MDG_LAUNCH_CHECK("diag_f64_kernel");
        tm.mark(3);
        rc = with_diag_words<L>(S, w, [&](auto nw) {
            with_group(w, [&](auto gs) {
                hipLaunchKernelGGL((reconstruct_kernel<L, decltype(nw)::value, decltype(gs)::value, long long>),
                                   dim3(blocks_for<int>(mn, 256)), dim3(256), 0, s, m, n, S, w,
                                   (const long long *)D, E, F, C, ldc);
            });
        });
        if (rc) return rc;
    }
    MDG_LAUNCH_CHECK("split gemm");
    tm.mark(4);
    float stage[4] = {0.f, 0.f, 0.f, 0.f};
    if (tm.on) MDG_HIP_CHECK(tm.elapsed(stage));
    if (info) {
        info->method = p.method;
        info->slices = p.S;
        info->slice_bits = p.w;
        info->diagonals = p.S;
        info->pair_gemms = (long long)p.S * (p.S + 1) / 2;
        info->total_bits = p.W;
        info->spread_a = p.spread_a;
        info->spread_b = p.spread_b;
        info->flushed = p.wide ? 1 : 0;
        info->backend = p.method == MDG_METHOD_INT8_OZAKI   ? (p.backend == I8E_LIB ? lib_engine_used : p.backend)
                        : p.method == MDG_METHOD_BF16_OZAKI ? p.backend
                                                            : 2;
        for (int i = 0; i < 4; ++i) info->stage_ms[i] = stage[i];
    }
    return MDG_SUCCESS;
}

int residue_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                 double *C, long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info,
                 hipStream_t s);

static int split_gemm_slices(int L, int m, int n, int k, const double *A, long long lda, const double *B,
                             long long ldb, double *C, long long ldc, int method, int target_bits, int fixed_spread,
                             mdg_gemm_info *info, hipStream_t s);

int split_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb, double *C,
               long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info, hipStream_t s)
{
    g_last_stats = nullptr;
    int rc;
    if (method == MDG_METHOD_INT8_RESIDUE || method == MDG_METHOD_F64_RESIDUE) {
        rc = residue_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, method, target_bits, fixed_spread, info, s);
    } else {
        rc = split_gemm_slices(L, m, n, k, A, lda, B, ldb, C, ldc, method, target_bits, fixed_spread, info, s);
    }
    if (rc || fixed_spread < 0) return rc;
    MDG_REQUIRE(g_last_stats, MDG_ERROR_HIP, "internal: no scan statistics for the guard");
    const long long nel = (long long)m * n * L;
    hipLaunchKernelGGL(guard_kernel, dim3((unsigned)std::min<long long>(1024, (nel + 255) / 256)), dim3(256), 0, s,
                       (const int *)g_last_stats, fixed_spread, m, n, C, ldc, L, g_guard_flag);
    MDG_LAUNCH_CHECK("guard_kernel");
    return MDG_SUCCESS;
}

static int split_gemm_slices(int L, int m, int n, int k, const double *A, long long lda, const double *B,
                             long long ldb, double *C, long long ldc, int method, int target_bits, int fixed_spread,
                             mdg_gemm_info *info, hipStream_t s)
{
    Plan p{};
    p.method = method;
    p.L = L;
    p.m = m;
    p.n = n;
    p.k = k;
    if (target_bits <= 0) target_bits = 53 * L;
    switch (L) {
    case 2: return run_split_gemm<2>(p, target_bits, fixed_spread, A, lda, B, ldb, C, ldc, s, info);
    case 4: return run_split_gemm<4>(p, target_bits, fixed_spread, A, lda, B, ldb, C, ldc, s, info);
    case 8: return run_split_gemm<8>(p, target_bits, fixed_spread, A, lda, B, ldb, C, ldc, s, info);
    case 16: return run_split_gemm<16>(p, target_bits, fixed_spread, A, lda, B, ldb, C, ldc, s, info);
    }
    set_error("unsupported limb count L=%d", L);
    return MDG_ERROR_INVALID_VALUE;
}

// ------------------------------------------- reference split-path pieces --
// The generic FP64 TN kernel serves the reference's accumulate GEMM
// (C++/TCtest.cu:21-46) and the limb convolution (Aq * bigB2(Bq)).

int f64_gemm_i32(const int32_t *A, long long lda, const int32_t *Bt, long long ldb, int m, int n, int depth, int flushk,
                 long long *D, hipStream_t s)
{
    const int tiles_m = (m + 127) / 128, tiles_n = (n + 127) / 128;
    dim3 grid(tiles_m * tiles_n, 1, 1);
    // one "diagonal" (S = 1) of the given depth, INT64 output plane of row stride n
    hipLaunchKernelGGL((diag_f64_kernel<int32_t, 2>), grid, dim3(512), 0, s, A, Bt, 1, depth, lda, ldb, m, n, tiles_m,
                       tiles_n, flushk, (void *)D, (long long)n);
    MDG_LAUNCH_CHECK("diag_f64_kernel (residue plane)");
    return MDG_SUCCESS;
}

int dgemm_acc(int M, int N, int K, const double *A, const double *B, double *C, hipStream_t s)
{
    const int tiles_m = (M + 127) / 128, tiles_n = (N + 127) / 128;
    dim3 grid(tiles_m * tiles_n, 1, 1);
    // one "diagonal" (S = 1) of depth K; B is column-major = K-contiguous rows
    hipLaunchKernelGGL((diag_f64_kernel<double, 1>), grid, dim3(512), 0, s, A, B, 1, K, (long long)K, (long long)K, M,
                       N, tiles_m, tiles_n, 1, (void *)C, (long long)N);
    MDG_LAUNCH_CHECK("diag_f64_kernel (dgemm_acc)");
    return MDG_SUCCESS;
}

// interleaved parts Xq[r][kk][t] (rows x k x pp) -> planes P[r][t*k + kk],
// reversed (t -> pp-1-t) when `rev`; transposed source (Bq[kk][c][t]) when `trans`
__global__ void parts_to_planes_kernel(int rows, int k, int pp, const double *__restrict__ Xq, long long row_stride,
                                       long long k_stride, int rev, double *__restrict__ P)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)rows * k * pp) return;
    const long long r = idx / ((long long)k * pp);
    const long long rem = idx - r * k * pp;
    const int t = (int)(rem / k);
    const int kk = (int)(rem - (long long)t * k);
    const int src_t = rev ? (pp - 1 - t) : t;
    P[r * ((long long)pp * k) + (long long)t * k + kk] = Xq[r * row_stride + kk * k_stride + src_t];
}

// D planes (pp x m x n) -> C1q[r][c][i]
__global__ void planes_to_parts_kernel(int m, int n, int pp, const double *__restrict__ D, double *__restrict__ C1q)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n * pp) return;
    const int i = (int)(idx % pp);
    const long long rc = idx / pp;
    C1q[idx] = D[(long long)i * m * n + rc];
}

// squeeze D planes directly: y[r][c] = fold_i inc_d(y, D_i[r][c])
template <int L>
__global__ __launch_bounds__(256) void squeeze_planes_kernel(long long nent, int pp, const double *__restrict__ D,
                                                             double *__restrict__ Y)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nent) return;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    for (int i = 0; i < pp; ++i) md_inc_d<L>(y, D[(long long)i * nent + idx]);
#pragma unroll
    for (int l = 0; l < L; ++l) Y[idx * L + l] = y[l];
}

// C1q (or planes) = Aq * bigB2(Bq) for m x k and k x n matrices of pp parts.
static int conv_planes(int m, int n, int k, int pp, const double *Aq, const double *Bq, double *D, hipStream_t s)
{
    double *Ap = (double *)scratch(s, sizeof(double) * (size_t)m * pp * k, 2);
    double *Bp = (double *)scratch(s, sizeof(double) * (size_t)n * pp * k, 3);
    MDG_REQUIRE(Ap && Bp, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    const long long na = (long long)m * k * pp, nb = (long long)n * k * pp;
    hipLaunchKernelGGL(parts_to_planes_kernel, dim3(blocks_for<int>(na, 256)), dim3(256), 0, s, m, k, pp, Aq,
                       (long long)k * pp, (long long)pp, 0, Ap);
    // Bq[kk][c][t]: "row" c has stride pp, kk has stride n*pp
    hipLaunchKernelGGL(parts_to_planes_kernel, dim3(blocks_for<int>(nb, 256)), dim3(256), 0, s, n, k, pp, Bq,
                       (long long)pp, (long long)n * pp, 1, Bp);
    const int tiles_m = (m + 127) / 128, tiles_n = (n + 127) / 128;
    dim3 grid(tiles_m * tiles_n, 1, (unsigned)pp);
    hipLaunchKernelGGL((diag_f64_kernel<double, 0>), grid, dim3(512), 0, s, Ap, Bp, pp, k, (long long)pp * k,
                       (long long)pp * k, m, n, tiles_m, tiles_n, 1, (void *)D, (long long)n);
    MDG_LAUNCH_CHECK("diag_f64_kernel (conv)");
    return MDG_SUCCESS;
}

int conv_gemm(int m, int n, int k, int pp, const double *Aq, const double *Bq, double *C1q, hipStream_t s)
{
    double *D = (double *)scratch(s, sizeof(double) * (size_t)pp * m * n, 4);
    MDG_REQUIRE(D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    int rc = conv_planes(m, n, k, pp, Aq, Bq, D, s);
    if (rc) return rc;
    const long long nel = (long long)m * n * pp;
    hipLaunchKernelGGL(planes_to_parts_kernel, dim3(blocks_for<int>(nel, 256)), dim3(256), 0, s, m, n, pp, D, C1q);
    MDG_LAUNCH_CHECK("planes_to_parts_kernel");
    return MDG_SUCCESS;
}

int conv_squeeze(int L, int m, int n, int k, int pp, const double *Aq, const double *Bq, double *C, hipStream_t s)
{
    double *D = (double *)scratch(s, sizeof(double) * (size_t)pp * m * n, 4);
    MDG_REQUIRE(D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    int rc = conv_planes(m, n, k, pp, Aq, Bq, D, s);
    if (rc) return rc;
    const long long nent = (long long)m * n;
    const dim3 grid(blocks_for<int>(nent, 256));
    switch (L) {
    case 2: hipLaunchKernelGGL(squeeze_planes_kernel<2>, grid, dim3(256), 0, s, nent, pp, D, C); break;
    case 4: hipLaunchKernelGGL(squeeze_planes_kernel<4>, grid, dim3(256), 0, s, nent, pp, D, C); break;
    case 8: hipLaunchKernelGGL(squeeze_planes_kernel<8>, grid, dim3(256), 0, s, nent, pp, D, C); break;
    case 16: hipLaunchKernelGGL(squeeze_planes_kernel<16>, grid, dim3(256), 0, s, nent, pp, D, C); break;
    default: set_error("unsupported limb count L=%d", L); return MDG_ERROR_INVALID_VALUE;
    }
    MDG_LAUNCH_CHECK("squeeze_planes_kernel");
    return MDG_SUCCESS;
}

}  // namespace mdg
