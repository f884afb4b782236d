// i8tile.hpp -- one TM x TM INT32 tile (TM = 128 or 64) of an INT8 x INT8
// product on the MFMA cores, shared by the fused batch pipelines (slice
// diagonals and residue planes).
//
// The tile is A (TM rows, row stride lda bytes) times B^T (TM rows, row
// stride ldb bytes) over a depth Kd that is a multiple of 128; every row of
// both operands must be readable up to Kd (the pipelines pad rows to 128 and
// depth with zero digits).  4 waves in a 2 x 2 grid, each TM/2 x TM/2 of
// v_mfma_i32_16x16x64_i8; the k-blocks are staged in LDS (rows padded by 16
// bytes for conflict-free 16-byte fragment reads) and the next RING
// k-blocks are fetched into registers while the current one is multiplied.
// INT32 sums
// are exact while Kd * 2^14 < 2^31.  The 64-row tile does a quarter of the
// work per workgroup: less padding on small problems, four times the
// workgroups.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdg {

typedef int i4t __attribute__((ext_vector_type(4)));

constexpr int I8T_BK = 128, I8T_LDR = I8T_BK + 16;

// 16-byte load from global memory.  The pipelines take their operand
// pointers from device tables, where the compiler cannot see the address
// space and would emit flat loads; a flat load also counts on the LDS wait
// counter, so every wait for an LDS fragment would wait for the prefetches
// in flight as well.
__device__ __forceinline__ i4t i8t_gload(const int8_t *p)
{
    return *(const __attribute__((address_space(1))) i4t *)p;
}

template <int TM>
struct I8Tile {
    static constexpr int IM = TM / 32;   // 16-row fragments per wave and dimension
    static constexpr int FP = TM / 32;   // 32-row fetch passes per operand
    // k-blocks of global loads in flight.  Two at TM = 128 keep its
    // occupancy (2 waves per SIMD) and cut the slice batch by 5%; at TM = 64
    // a deeper ring costs occupancy (6 -> 4 waves per SIMD) and was slower
    static constexpr int RING = TM == 64 ? 1 : 2;
};

// As, Bs: LDS buffers of TM * I8T_LDR bytes each (16-byte aligned).
// R register sets hold the global loads of the next R k-blocks, so a tile
// of up to R k-blocks pays one fetch latency instead of one per block (the
// pipelines' tiles are a few blocks deep).  Slots whose block lies past the
// end re-read the last block (in bounds, cached) instead of branching
// around the loads, and their LDS pass is skipped.
template <int TM, int R = 1>
__device__ __forceinline__ void i8_tile(const int8_t *__restrict__ Ab, long long lda, const int8_t *__restrict__ Bb,
                                        long long ldb, int Kd, int8_t *As, int8_t *Bs,
                                        i4t (&acc)[I8Tile<TM>::IM][I8Tile<TM>::IM])
{
    constexpr int LDR = I8T_LDR, BK = I8T_BK, IM = I8Tile<TM>::IM, FP = I8Tile<TM>::FP, HW = TM / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 15, lk = (lane >> 4) * 16;
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < IM; ++j) acc[i][j] = (i4t){0, 0, 0, 0};
    const int frow = tid >> 3, fseg = (tid & 7) * 16;
    const long long fa = (long long)frow * lda + fseg, fsa = 32 * lda;
    const long long fb = (long long)frow * ldb + fseg, fsb = 32 * ldb;
    const int soff = frow * LDR + fseg, sstep = 32 * LDR;
    const int nkb = Kd / BK, last = nkb - 1;
    i4t ra[R][FP], rb[R][FP];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const long long ko = (long long)min(r, last) * BK;
#pragma unroll
        for (int f = 0; f < FP; ++f) {
            ra[r][f] = i8t_gload(Ab + fa + f * fsa + ko);
            rb[r][f] = i8t_gload(Bb + fb + f * fsb + ko);
        }
    }
    for (int kb0 = 0; kb0 < nkb; kb0 += R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int kb = kb0 + r;
            const bool live = kb < nkb;
            if (live) {
#pragma unroll
                for (int f = 0; f < FP; ++f) {
                    *(i4t *)&As[soff + f * sstep] = ra[r][f];
                    *(i4t *)&Bs[soff + f * sstep] = rb[r][f];
                }
            }
            __syncthreads();
            {
                const long long ko = (long long)min(kb + R, last) * BK;
#pragma unroll
                for (int f = 0; f < FP; ++f) {
                    ra[r][f] = i8t_gload(Ab + fa + f * fsa + ko);
                    rb[r][f] = i8t_gload(Bb + fb + f * fsb + ko);
                }
            }
            if (live) {
#pragma unroll
                for (int ks = 0; ks < BK; ks += 64) {
                    i4t a[IM], b[IM];
#pragma unroll
                    for (int i = 0; i < IM; ++i) a[i] = *(const i4t *)&As[(wm * HW + i * 16 + lrow) * LDR + ks + lk];
#pragma unroll
                    for (int j = 0; j < IM; ++j) b[j] = *(const i4t *)&Bs[(wn * HW + j * 16 + lrow) * LDR + ks + lk];
#pragma unroll
                    for (int i = 0; i < IM; ++i)
#pragma unroll
                        for (int j = 0; j < IM; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
                }
            }
            __syncthreads();
        }
    }
}

// calls put(row, col, value) for the tile's in-range entries, rows < m,
// cols < n, with the tile at (row0, col0)
template <int TM, typename Put>
__device__ __forceinline__ void i8_tile_each(const i4t (&acc)[I8Tile<TM>::IM][I8Tile<TM>::IM], long long row0,
                                             long long col0, int m, int n, Put put)
{
    constexpr int IM = I8Tile<TM>::IM, HW = TM / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < IM; ++j) {
            const long long col = col0 + wn * HW + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long row = row0 + wm * HW + i * 16 + (lane >> 4) * 4 + v;
                if (row < m && col < n) put(row, col, acc[i][j][v]);
            }
        }
}

}  // namespace mdg
