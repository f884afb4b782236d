// blkslice.hpp -- slicing into the blocked layout of the band kernel, one
// tile of rows x one k-block per workgroup.
//
// The band kernel reads its slices as 16 KiB blocks (i8band.hpp): block
// (tile, kb, t) holds slice t of 128 rows x 128 k-bytes, each row's 16-byte
// chunks permuted by the row swizzle.  A thread per entry that stores its S
// digits one byte at a time issues S byte stores per entry; here a
// workgroup converts RT rows x one k-block of entries (EB = 128 / sizeof(T)
// per row, RT = 256 / EB: one entry per thread), gathers their digits in LDS
// as RT x S natural-order rows of 128 bytes, and writes every (row, slice)
// out as eight 16-byte chunks, already swizzled: 16 times fewer store
// instructions for INT8, and full 128-byte lines per row.  Entries are read
// k-fastest for A (rows contiguous along k) and rows-fastest for B (its rows
// of the k x n operand contiguous along n, RT entries per 128-byte line).
// Padding entries (beyond rows or k) and rows of scale ZERO (all-zero
// rows/columns) get zero digits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "i8band.hpp"
#include "slices.hpp"

namespace mdg {

template <typename T>
__device__ __forceinline__ T blk_digit(int d);
template <>
__device__ __forceinline__ int8_t blk_digit<int8_t>(int d) { return (int8_t)d; }
template <>
__device__ __forceinline__ __bf16 blk_digit<__bf16>(int d) { return (__bf16)(float)d; }

// LDS staging per workgroup: RT rows of S slices of 128 bytes, each row
// padded by 16 bytes so the rows of one k-offset fall in different banks
constexpr int BLK_STAGE_MAX = 33024;
__host__ __device__ __forceinline__ int blk_row_bytes(int S) { return S * 128 + 16; }

// rows per workgroup tile for digits of `digit_bytes` bytes: one entry per
// thread (256 / EB rows of EB entries), fewer rows when the staging of deep
// slice stacks would exceed BLK_STAGE_MAX
__host__ __device__ __forceinline__ int blk_rows_per_tile(int S, int digit_bytes = 1)
{
    int rt = 256 / (BND_BK / digit_bytes);
    while (rt > 1 && rt * blk_row_bytes(S) > BLK_STAGE_MAX) rt >>= 1;
    return rt;
}

// dynamic LDS bytes of a slicing workgroup
__host__ __device__ __forceinline__ int blk_stage_bytes(int S, int digit_bytes = 1)
{
    return blk_rows_per_tile(S, digit_bytes) * blk_row_bytes(S);
}

// Fixed-point words a slicing of S digits of w bits needs, |N| < 2^(S w - 2):
// the smallest of three register-array sizes per limb count (the largest
// FixedWords<L>::NW), so the conversion's word loop runs no further than N
// reaches.  fn(std::integral_constant<int, NW>).
template <int L, typename Fn>
__host__ __device__ __forceinline__ void with_slice_words(int S, int w, Fn fn)
{
    constexpr int NWMAX = FixedWords<L>::NW;
    constexpr int NW1 = (NWMAX + 1) / 2, NW2 = (2 * NWMAX + 2) / 3;
    const int need = (S * w + 31) / 32;
    if (need <= NW1)
        fn(std::integral_constant<int, NW1>{});
    else if (need <= NW2)
        fn(std::integral_constant<int, NW2>{});
    else
        fn(std::integral_constant<int, NWMAX>{});
}

// Slices of rows r0 .. r0+RT-1 (RT = blk_rows_per_tile(S, sizeof(T))),
// k-block kb, of an operand of `rows` valid rows and depth k (padded to
// kp).  TRANS: X is
// B (k x n, row stride ldx) and "row" r is column r of B; else X is A (rows
// x k).  E[r]: the row's scale exponent, zero_scale for an all-zero row.
// Entry with fixed-point N = floor(x 2^(S w - 2 - E)) gets its S balanced
// digits (slices.hpp) in NW words (with_slice_words).  All 256 threads of
// the workgroup call this; stage: blk_stage_bytes(S, sizeof(T)) bytes of
// LDS, 16-byte aligned.  KEEP_PAD: the padding (rows from `rows` on, and
// 16-byte chunks wholly past k) already holds zero digits and is not
// written again.
template <int L, bool TRANS, typename T, int NW, bool KEEP_PAD = false>
__device__ __forceinline__ void slice_tile(int rows, int k, int kp, const double *__restrict__ X, long long ldx,
                                           const int *__restrict__ E, int zero_scale, int S, int w, int kb, int r0,
                                           int8_t *__restrict__ out, int8_t *stage)
{
    static_assert(NW <= FixedWords<L>::NW, "fixed-point words");
    constexpr int EB = BND_BK / (int)sizeof(T);
    const int RT = blk_rows_per_tile(S, (int)sizeof(T));
    const int rb = blk_row_bytes(S);
    for (int e = threadIdx.x; e < RT * EB; e += 256) {
        const int rr = TRANS ? e % RT : e / EB;
        const int kl = TRANS ? e / RT : e % EB;
        const int r = r0 + rr, kk = kb * EB + kl;
        T *o = (T *)(stage + rr * rb) + kl;   // slice t at o + t EB
        const int ex = (r < rows) ? E[r] : zero_scale;
        if (r >= rows || kk >= k || ex == zero_scale) {
            for (int t = 0; t < S; ++t) o[t * EB] = blk_digit<T>(0);
            continue;
        }
        double x[L];
#pragma unroll
        for (int l = 0; l < L; ++l)
            x[l] = TRANS ? X[((long long)kk * ldx + r) * L + l] : X[((long long)r * ldx + kk) * L + l];
        uint32_t N[NW];
        to_fixed<L, NW>(x, S * w - 2 - ex, N);
        auto emit = [&](int t, int d) { o[t * EB] = blk_digit<T>(d); };
        if (w == 8)
            fixed_digits8<NW>(N, S, emit);
        else
            fixed_digits<NW>(N, S, w, emit);
    }
    __syncthreads();
    // RT x S rows of 128 bytes, eight 16-byte chunks each; chunk ch of row r
    // lands at chunk (ch ^ swz(r)) of its block row
    const int nkb = kp / EB;
    const int kbytes = (k - kb * EB) * (int)sizeof(T);   // valid bytes of each row of this k-block
    for (int c = threadIdx.x; c < RT * S * 8; c += 256) {
        const int ch = c & 7, rs = c >> 3;
        const int rr = rs / S, t = rs - rr * S;
        const int r = r0 + rr;
        if (KEEP_PAD && (r >= rows || ch * 16 >= kbytes)) continue;
        const int4 v = *(const int4 *)(stage + rr * rb + t * 128 + ch * 16);
        const int sw = (r & 7) ^ ((r >> 3) & 1);
        *(int4 *)(out + bnd_block(r >> 7, kb, t, S, nkb) + (r & 127) * BND_BK + (((ch ^ sw) & 7) << 4)) = v;
    }
}

}  // namespace mdg
