// i8tile.hpp -- one 128 x 128 INT32 tile of an INT8 x INT8 product on the
// MFMA cores, shared by the fused batch pipelines (slice diagonals and
// residue planes).
//
// The tile is A (128 rows, row stride lda bytes) times B^T (128 rows, row
// stride ldb bytes) over a depth Kd that is a multiple of 128; every row of
// both operands must be readable up to Kd (the pipelines pad rows and depth
// with zero digits).  4 waves, each 64 x 64 of v_mfma_i32_16x16x64_i8; the
// k-blocks are staged in LDS (rows padded by 16 bytes for conflict-free
// 16-byte fragment reads) and the next k-block is fetched into registers
// while the current one is multiplied.  INT32 sums are exact while
// Kd * 2^14 < 2^31.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdg {

typedef int i4t __attribute__((ext_vector_type(4)));

constexpr int I8T_BM = 128, I8T_BK = 128, I8T_LDR = I8T_BK + 16;

// As, Bs: LDS buffers of I8T_BM * I8T_LDR bytes each (16-byte aligned)
__device__ __forceinline__ void i8_tile_128(const int8_t *__restrict__ Ab, long long lda,
                                            const int8_t *__restrict__ Bb, long long ldb, int Kd, int8_t *As,
                                            int8_t *Bs, i4t (&acc)[4][4])
{
    constexpr int LDR = I8T_LDR, BK = I8T_BK;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = lane & 15, lk = (lane >> 4) * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (i4t){0, 0, 0, 0};
    const int frow = tid >> 3, fseg = (tid & 7) * 16;
    const long long fa = (long long)frow * lda + fseg, fsa = 32 * lda;
    const long long fb = (long long)frow * ldb + fseg, fsb = 32 * ldb;
    const int soff = frow * LDR + fseg, sstep = 32 * LDR;
    i4t ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define MDG_I8T_FETCH(kb)                                 \
    do {                                                  \
        ra0 = *(const i4t *)(Ab + fa + (kb));             \
        ra1 = *(const i4t *)(Ab + fa + fsa + (kb));       \
        ra2 = *(const i4t *)(Ab + fa + 2 * fsa + (kb));   \
        ra3 = *(const i4t *)(Ab + fa + 3 * fsa + (kb));   \
        rb0 = *(const i4t *)(Bb + fb + (kb));             \
        rb1 = *(const i4t *)(Bb + fb + fsb + (kb));       \
        rb2 = *(const i4t *)(Bb + fb + 2 * fsb + (kb));   \
        rb3 = *(const i4t *)(Bb + fb + 3 * fsb + (kb));   \
    } while (0)
    MDG_I8T_FETCH(0);
    for (int kb = 0; kb < Kd; kb += BK) {
        *(i4t *)&As[soff] = ra0;
        *(i4t *)&As[soff + sstep] = ra1;
        *(i4t *)&As[soff + 2 * sstep] = ra2;
        *(i4t *)&As[soff + 3 * sstep] = ra3;
        *(i4t *)&Bs[soff] = rb0;
        *(i4t *)&Bs[soff + sstep] = rb1;
        *(i4t *)&Bs[soff + 2 * sstep] = rb2;
        *(i4t *)&Bs[soff + 3 * sstep] = rb3;
        __syncthreads();
        if (kb + BK < Kd) MDG_I8T_FETCH(kb + BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 64) {
            i4t a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = *(const i4t *)&As[(wm * 64 + i * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = *(const i4t *)&Bs[(wn * 64 + j * 16 + lrow) * LDR + ks + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#undef MDG_I8T_FETCH
}

// the tile's in-range entries into plane[row * ldp + col], rows < m, cols < n
__device__ __forceinline__ void i8_tile_store(const i4t (&acc)[4][4], int *plane, long long ldp, long long row0,
                                              long long col0, int m, int n)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long col = col0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long row = row0 + wm * 64 + i * 16 + (lane >> 4) * 4 + v;
                if (row < m && col < n) plane[row * ldp + col] = acc[i][j][v];
            }
        }
}

}  // namespace mdg
