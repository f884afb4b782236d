// i8plain.hpp -- strided-batched INT8 x INT8 -> INT32 GEMM on MFMA
//
//   D_b[r * ldd + c] = sum_{kk < depth} A_b[r * lda + kk] * Bt_b[c * ldb + kk]
//
// for r < m, c < n, with A_b = A + b * sa and Bt_b = Bt + b * sb: both
// operands row-major and contiguous along the depth (the residue planes of
// residue_gemm.hip).  This is the GEMM of one modulus (or one digit pair of
// a wide modulus) of the residue method; the sums are exact INT32 while
// depth * 2^14 < 2^31.
//
// Workgroup tile 256 x 256, 4 waves in 2 x 2, each 128 x 128 = 4 x 4 tiles
// of v_mfma_i32_32x32x32_i8 (256 INT32 sums per lane in the accumulation
// registers).  A k-block of 128 bytes is a stage of four 16 KiB LDS blocks
// (A rows 0-127, 128-255, Bt rows 0-127, 128-255) in the layout of the band
// kernel (i8band.hpp: 128-byte rows, 16-byte segments swizzled, conflict-free
// fragment reads).  The copies go straight from global memory to LDS
// (global_load_lds, 16 bytes per lane); the swizzle is applied on the
// source side: lane l of a 1 KiB copy lands at LDS segment l % 8 of row
// l / 8 and so reads the k-bytes of segment (l % 8) ^ swz(row).  Two stages:
// the copies of k-block s+2 go out at the barrier of step s, into the stage
// step s has finished reading.  Per k-step a wave reads 8 fragments for 16
// MFMAs.
//
// A plain GEMM moves 256 operand bytes per MFMA through the copies (the band
// kernel, whose A blocks serve four diagonals, half that); that traffic, not
// the MFMAs, sets its pace: about 0.33 of the INT8 peak on 4096^3 planes
// with random digits, against 0.54 with the copies left out.  Deeper rings
// (five stages of 64-byte steps) and copies staged through registers
// measured no better.
#pragma once

#include "i8band.hpp"

namespace mdg {

constexpr int PLN_TM = 256;
constexpr int PLN_BK = BND_BK;   // bytes of depth per step

template <int dummy = 0>
__global__ __launch_bounds__(256) void plain_i8_kernel(const int8_t *__restrict__ A, long long lda, long long sa, int ma,
                                                       const int8_t *__restrict__ Bt, long long ldb, long long sb,
                                                       int nbr, int m, int n, int depth, int *__restrict__ D,
                                                       long long ldd, long long sd, int tiles_m, int tiles_n)
{
    constexpr int STG = 4 * BND_BLK;   // A0 A1 B0 B1
    __shared__ __attribute__((aligned(1024))) int8_t lds[2 * STG];
    const int bat = blockIdx.y;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const int nkb = depth / BND_BK;

    // copies: wave w fills block w of a stage: rows 128 w .. of the A tile
    // (w < 2) or rows 128 (w - 2) .. of the Bt tile; rows past the operand's
    // readable rows (ma of A, nbr of Bt) reread its last row (their sums are
    // never stored).  Copy i moves rows 8 i .. 8 i + 7.
    const bool isA = wave < 2;
    const int8_t *src = isA ? A + bat * sa : Bt + bat * sb;
    const long long ld = isA ? lda : ldb;
    const int rlim = (isA ? ma : nbr) - 1;
    const int r0 = (isA ? tm : tn) * PLN_TM + (wave & 1) * BND_TM;
    const int8_t *p[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = 8 * i + (lane >> 3);
        const int sw = (row & 7) ^ ((row >> 3) & 1);
        p[i] = src + (long long)min(r0 + row, rlim) * ld + 16 * ((lane & 7) ^ sw);
    }
    int8_t *ldsw = lds + wave * BND_BLK;
    auto issue = [&](int kb, int stage) {
        kb = min(kb, nkb - 1);   // past the end: reload the last k-block (nobody reads it)
        const long long ko = (long long)kb * BND_BK;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            auto *d = (__attribute__((address_space(3))) void *)(ldsw + stage * STG + i * 1024);
            __builtin_amdgcn_global_load_lds((const void *)(p[i] + ko), d, 16, 0, 0);
        }
    };

    // fragments: lane l reads row l % 32 of a 32-row group, k-bytes
    // 16 (l / 32) .. + 16 of each 32-byte k-step
    const int lrow = lane & 31, hk = lane >> 5;
    const int sw = (lrow & 7) ^ ((lrow >> 3) & 1);
    int koff[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) koff[ks] = (((2 * ks + hk) ^ sw) & 7) << 4;
    const int8_t *fragA = lds + wm * BND_BLK + lrow * BND_BK;
    const int8_t *fragB = lds + (2 + wn) * BND_BLK + lrow * BND_BK;

    bnd_i16 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

    bnd_i4 a[2][4], b[2][4];
    int stage = 0;
    auto fetch = [&](int buf, int ks) {
        const int so = stage * STG + koff[ks];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[buf][i] = *(const bnd_i4 *)(fragA + so + i * 4096);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[buf][j] = *(const bnd_i4 *)(fragB + so + j * 4096);
    };
    auto mfma = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[buf][i], b[buf][j], acc[i][j], 0, 0, 0);
    };

    // prologue: k-blocks 0 and 1 into stages 0 and 1
    issue(0, 0);
    issue(1, 1);
    bnd_wait_vm<16>();
    __builtin_amdgcn_s_barrier();
    fetch(0, 0);
    // step s: k-steps 0-2 read the next k-step's fragments under the MFMAs
    // of this one; then wait for k-block s+1, barrier (every wave is done
    // with stage s), send k-block s+2 into stage s, read k-block s+1's first
    // fragments, and run k-step 3's MFMAs under the copies and reads
    for (int s = 0; s < nkb; ++s) {
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            fetch((ks & 1) ^ 1, ks + 1);
            mfma(ks & 1);
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // LDS read
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        bnd_wait_vm<0>();
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(s + 2, stage);
        stage ^= 1;
        fetch(0, 0);
        mfma(1);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // LDS read
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x020, 2, 0);   // global->LDS copies
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    // no global->LDS copy may still be in flight when the workgroup ends
    bnd_wait_vm<0>();

    // acc element v of a 32 x 32 tile: row 8 (v / 4) + 4 (lane / 32) + v % 4,
    // column lane % 32
    int *Db = D + bat * sd;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long col = (long long)tn * PLN_TM + wn * 128 + j * 32 + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long row = (long long)tm * PLN_TM + wm * 128 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (row < m && col < n) Db[row * ldd + col] = acc[i][j][v];
            }
        }
}

}  // namespace mdg
