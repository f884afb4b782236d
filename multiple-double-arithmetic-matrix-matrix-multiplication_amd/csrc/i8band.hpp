// i8band.hpp -- the INT8 slice-pair GEMM: bands of diagonals per workgroup.
//
// Diagonal g of the slice product is D_g = sum_{t+u=g} A_t * B_u (A_t, B_u
// the INT8 slices of A's rows and B's columns, K-contiguous).  Diagonal g on
// its own is a GEMM of depth (g+1) k that re-reads slices 0..g of both
// operands; a workgroup here computes a BAND of G consecutive diagonals
// g0 .. g0+G-1 of one 128 x 128 output tile at once.  Per k-block it streams
// A_t for t = 0, 1, ... and keeps a sliding window of the G slices
// B_{g0-t} .. B_{g0+G-1-t} in LDS: every A block meets G B blocks, and every
// B block is fetched once per k-block instead of once per diagonal, which
// halves the operand traffic per MFMA against a plain 256 x 256 tile.
//
// Operand layout ("blocked"): 128 rows x 128 k-bytes per 16 KiB block,
// block (tile, kb, slice) at ((tile * nkb + kb) * S + slice) * 16 KiB, so a
// workgroup's stream of slices of one k-block is contiguous.  Inside a block
// row r holds its 128 bytes as eight 16-byte chunks in the order
// chunk ^ swz(r), swz(r) = (r & 7) ^ ((r >> 3) & 1): a block is copied to LDS
// as-is by direct global->LDS loads (1 KiB contiguous per wave instruction)
// and the 16-byte MFMA fragment reads of 16 consecutive rows then hit 16
// distinct 4-bank groups (and 8 rows, 8 groups): no bank conflicts.
//
// Exactness: |d_t d_u| <= 2^14, so an INT32 sum of at most 131071 products
// is exact.  A diagonal of band q sums at most n_t = g0_q + G slice pairs
// per element of depth, so the band's k-blocks are cut into K-chunks of
// floor(cap / n_t) k-blocks (cap = floor(131071 / elements per k-block)),
// and each (band, chunk) writes one INT64 "band plane": its G diagonal sums
// combined exactly, P = sum_j D_{g0+j} 2^(8 (G-1-j)) (|P| < 2^56), half
// the bytes of one INT32 plane per diagonal (recon.hpp band_fold takes them
// apart again and adds the chunks).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <vector>

namespace mdg {

constexpr int BND_TM = 128;                 // tile rows = tile columns
constexpr int BND_BK = 128;                 // k-bytes per block
constexpr int BND_BLK = BND_TM * BND_BK;    // 16 KiB

// swizzled byte position of (row r, k-byte kk) inside a block
__host__ __device__ __forceinline__ int bnd_pos(int r, int kk)
{
    const int sw = (r & 7) ^ ((r >> 3) & 1);
    return r * BND_BK + ((((kk >> 4) ^ sw) & 7) << 4) + (kk & 15);
}

// byte offset of block (tile, kb, slice) of an operand of S slices and nkb
// k-blocks per slice
__host__ __device__ __forceinline__ long long bnd_block(int tile, int kb, int t, int S, int nkb)
{
    return (((long long)tile * nkb + kb) * S + t) * (long long)BND_BLK;
}

// One work item of the band kernel: diagonals g0 .. g0+gb-1 over k-blocks
// [kb0, kb1) (one K-chunk), every slice t of A with a pair in the band,
// t = 0 .. g0+gb-1.  g0 < 0 for the band of leftover diagonals
// (band_items): diagonals below 0 are empty.
struct BandItem {
    short g0, gb, kb0, kb1;
};

// products per INT32 accumulator of one K-chunk, over elements per k-block
constexpr int BND_CAP_I8 = 131071 / BND_BK;          // INT8: 128 per k-block
constexpr int BND_CAP_BF16 = 131071 / (BND_BK / 2);  // BF16: 64 per k-block

constexpr int BND_MAX_ITEMS = 240;

struct BandTable {
    int count;
    int pad;
    BandItem item[BND_MAX_ITEMS];
};

typedef int bnd_i4 __attribute__((ext_vector_type(4)));
typedef int bnd_i16 __attribute__((ext_vector_type(16)));
typedef float bnd_f16 __attribute__((ext_vector_type(16)));
typedef __bf16 bnd_bf8 __attribute__((ext_vector_type(8)));

// f(integral_constant<int, R>) for R in the sequence, in order
template <int... R, typename F>
__device__ __forceinline__ void bnd_unroll(std::integer_sequence<int, R...>, F &&f)
{
    (f(std::integral_constant<int, R>{}), ...);
}

// s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding: vmcnt[3:0] in bits 3:0,
// vmcnt[5:4] in bits 15:14, expcnt left at its maximum, lgkmcnt 0): every
// global->LDS copy but the last N instructions, and every LDS access, done
template <int N>
__device__ __forceinline__ void bnd_wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt out of range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0070);
}

// 4 KiB global -> LDS: four copies of 16 bytes per lane, each wave
// instruction filling 1 KiB at lds_wave + its immediate offset (the
// instruction offset advances the global and the LDS address alike, so the
// four share one address register and one M0 value)
__device__ __forceinline__ void bnd_dma4k(const int8_t *src, int8_t *lds_wave)
{
    auto *d = (__attribute__((address_space(3))) void *)lds_wave;
    __builtin_amdgcn_global_load_lds((const void *)src, d, 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void *)src, d, 16, 1024, 0);
    __builtin_amdgcn_global_load_lds((const void *)src, d, 16, 2048, 0);
    __builtin_amdgcn_global_load_lds((const void *)src, d, 16, 3072, 0);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Give every XCD a contiguous block of tiles and walk that block in
// groups of GROUP tile-rows, so the tiles resident on one XCD share A rows
// and B columns in its L2.
__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int &tm, int &tn)
{
    const int T = tiles_m * tiles_n;
    int lin = bid;
    if ((T & 7) == 0) lin = (bid & 7) * (T >> 3) + (bid >> 3);
    constexpr int GROUP = 8;
    const int per_group = GROUP * tiles_n;
    const int gidx = lin / per_group;
    const int first_m = gidx * GROUP;
    const int gsize = min(tiles_m - first_m, GROUP);
    const int in = lin - gidx * per_group;
    tm = first_m + in % gsize;
    tn = in / gsize;
}

// Band planes.  Band q = 0 .. nb-1, nb = ceil(S / G), holds diagonals g0_q
// .. g0_q + G - 1 with g0_q = S - G (nb - q) (band 0 starts below diagonal
// 0 when G does not divide S, band_items); its n_t = g0_q + G slices of A
// give K-chunks of floor(cap / n_t) k-blocks, ceil(nkb / that) of them;
// plane (q, c) is bnd_first_band_plane(q) + c.
__host__ __device__ __forceinline__ int bnd_band_kblocks(int S, int G, int cap, int q)
{
    const int nb = (S + G - 1) / G;
    const int nt = S - G * (nb - q) + G;
    return max(1, cap / nt);
}

__host__ __device__ __forceinline__ int bnd_band_chunks(int S, int G, int nkb, int cap, int q)
{
    const int kc = bnd_band_kblocks(S, G, cap, q);
    return (nkb + kc - 1) / kc;
}

__host__ __device__ __forceinline__ int bnd_first_band_plane(int S, int G, int nkb, int cap, int q)
{
    int p = 0;
    for (int h = 0; h < q; ++h) p += bnd_band_chunks(S, G, nkb, cap, h);
    return p;
}

__host__ __device__ __forceinline__ int bnd_band_planes(int S, int G, int nkb, int cap)
{
    return bnd_first_band_plane(S, G, nkb, cap, (S + G - 1) / G);
}

// the chunk count of every band, for the reconstruction (a kernel argument:
// no division per band on the device); nb <= ceil(255 / 3)
struct BandPlanes {
    int nb, total;
    unsigned short nch[96];
};

inline BandPlanes bnd_plane_table(int S, int G, int nkb, int cap)
{
    BandPlanes t{};
    t.nb = (S + G - 1) / G;
    for (int q = 0; q < t.nb && q < 96; ++q) {
        t.nch[q] = (unsigned short)bnd_band_chunks(S, G, nkb, cap, q);
        t.total += t.nch[q];
    }
    return t;
}

// Band kernel.  Grid: x = output tiles (tiles_m * tiles_n), y = band items.
// 4 waves in WM x (4 / WM), each a (128 / WM) x (128 WM / 4) piece of every
// diagonal of the band in v_mfma_i32_32x32x32_i8 tiles, G * 64 INT32 sums per
// lane in the accumulation registers.  WM = 1 (waves side by side, each 128
// rows x 32 columns) reads 4 A and G B fragments per 4 G MFMAs, against
// 2 + 2 G for WM = 2: fewer LDS reads, 1-2% faster at quad-double n = 4096.
//
// Step s of the workgroup: k-block kb = kb0 + s / nsk, slice t = t0 + s % nsk.
// The A stream holds A_t[kb] at position s; the B stream holds, per k-block,
// B_u for u = g0+gb-1-t0 down to max(0, g0-t1+1), and step s needs its
// positions s .. s+gb-1 (the window).  Position p of the A stream lives in
// LDS slot p % RA, of the B stream in slot p % RB; loads run PA steps
// ahead.  Each step: wait for this wave's loads of step s, barrier (every
// wave's loads of step s landed, every wave done with step s-1), issue the
// loads of step s+PA into the slots step s-1 freed, multiply.  A chunk that
// ends before the band's last diagonal (t1 <= g0) pads each k-block with
// gb-1 steps that only load B.
template <int G, int PA>
constexpr int bnd_lds_bytes()
{
    // slots: PA + 1 of the A stream, PA + G of the B stream, one block of zeros
    return (PA + 1 + PA + G + 1) * BND_BLK;
}

// One band item of one output tile: Atile / Btile = the tile's blocked
// slices (block (kb, slice) at (kb * S + slice) * 16 KiB), diagonals
// g0 .. g0+G-1, slices t of A in [t0, t1), k-blocks [kb0, kb1) of the nkb
// (K-chunk kb0 / bnd_band_kblocks of the band, for chunks of cap: see
// bnd_band_kblocks); the tile's rows start at row0 of the m x n band planes
// D (INT64), its columns at col0.  lds:
// bnd_lds_bytes<G, PA>() bytes.  IMV < IM (INT8) keeps only the first IMV 32-row MFMA tiles of
// every wave: the work on the zero padding of a tile whose rows from
// row0 + 32 IMV on are beyond m is left out.
template <int G, int PA, int WM, bool BF = false, int IMV = BND_TM / WM / 32>
__device__ __forceinline__ void band_run(int8_t *lds, const int8_t *__restrict__ Atile,
                                         const int8_t *__restrict__ Btile, int S, int nkb, int cap, int g0, int t0,
                                         int t1, int kb0, int kb1, long long row0, long long col0, int m, int n,
                                         long long *__restrict__ D)
{
    constexpr int RA = PA + 1, RB = PA + G;
    // WM x WN waves, each TMW x TNW of every diagonal: IM x IN MFMA tiles
    constexpr int WN = 4 / WM, TMW = BND_TM / WM, TNW = BND_TM / WN, IM = TMW / 32, IN = TNW / 32;
    static_assert(bnd_lds_bytes<G, PA>() <= 163840, "LDS ring exceeds 160 KiB");
    static_assert(IMV >= 1 && IMV <= IM && (!BF || IMV == IM), "row tiles: all of them for BF16");
    constexpr int gb = G;   // items of one launch have exactly G diagonals
    const int nA = t1 - t0;
    const int nsk = (t1 <= g0 + 1) ? nA + gb - 1 : g0 + gb - t0;
    const int uhi = g0 + gb - 1 - t0;
    const int nsteps = (kb1 - kb0) * nsk;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave / WN, wn = wave % WN;
    const int8_t *Arow = Atile + wave * 4096 + lane * 16;   // per-lane base of this wave's 4 KiB
    const int8_t *Brow = Btile + wave * 4096 + lane * 16;
    int8_t *ldsA = lds + wave * 4096;
    int8_t *ldsB = lds + RA * BND_BLK + wave * 4096;

    // issue cursors: A position pa = s + PA, B position pb = s + PA + gb - 1,
    // as (k-block, index within the k-block) and ring slot
    int a_kb = kb0, a_i = 0, a_slot = 0, b_kb = kb0, b_i = 0, b_slot = 0;
    auto issue_a = [&]() {
        int kb = a_kb, i = a_i;
        if (kb >= kb1) {   // past the end: reload the last block (in bounds, nobody reads it)
            kb = kb1 - 1;
            i = nsk - 1;
        }
        const int t = t0 + min(i, nA - 1);
        const int8_t *src = Arow + ((long long)kb * S + t) * BND_BLK;
        bnd_dma4k(src, ldsA + a_slot * BND_BLK);
        if (++a_i == nsk) {
            a_i = 0;
            ++a_kb;
        }
        if (++a_slot == RA) a_slot = 0;
    };
    auto issue_b = [&]() {
        int kb = b_kb, i = b_i;
        if (kb >= kb1) {
            kb = kb1 - 1;
            i = nsk - 1;
        }
        const int8_t *src = Brow + ((long long)kb * S + (uhi - i)) * BND_BLK;
        bnd_dma4k(src, ldsB + b_slot * BND_BLK);
        if (++b_i == nsk) {
            b_i = 0;
            ++b_kb;
        }
        if (++b_slot == RB) b_slot = 0;
    };

    // prologue: B positions 0 .. gb-2, then the pairs (A p, B p+gb-1), p < PA
    for (int p = 0; p < gb - 1; ++p) issue_b();
#pragma unroll
    for (int p = 0; p < PA; ++p) {
        issue_a();
        issue_b();
    }

    // fragment addresses: lane l reads row (or column) l % 32 of a 32-row
    // group and k-bytes 16 (l / 32) .. +16 of each 32-byte k-step
    const int lrow = lane & 31, hk = lane >> 5;
    const int sw = (lrow & 7) ^ ((lrow >> 3) & 1);
    int koff[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) koff[ks] = (((2 * ks + hk) ^ sw) & 7) << 4;
    const int8_t *fragA = lds + (wm * TMW + lrow) * BND_BK;
    const int8_t *fragB = lds + RA * BND_BLK + (wn * TNW + lrow) * BND_BK;
    {
        // the zero block (slot RB of the B ring); the first barrier of the
        // main loop orders these writes before every read
        bnd_i4 *z = (bnd_i4 *)(lds + (RA + RB) * BND_BLK);
#pragma unroll
        for (int i = 0; i < BND_BLK / (16 * 256); ++i) z[threadIdx.x + 256 * i] = (bnd_i4){0, 0, 0, 0};
    }

    // BF (BF16 slices): v_mfma_f32_32x32x16_bf16 reads its two 16-byte
    // fragments with the byte layout of v_mfma_i32_32x32x32_i8 (8 bf16 of a
    // 16-element k-step per lane), so blocks, ring and fragment reads are the
    // same; the FP32 sums of each tile are moved into INT32 sums (tot) while
    // they are still exact integers
    using AccT = std::conditional_t<BF, bnd_f16, bnd_i16>;
    AccT acc[G][IM][IN];
    bnd_i16 tot[BF ? G : 1][IM][IN];
    if constexpr (BF) {
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int jj = 0; jj < IN; ++jj)
#pragma unroll
                    for (int v = 0; v < 16; ++v) tot[j][i][jj][v] = 0;
    }
    // BF16 tiles are moved into tot KT at a time, one group per moving step
    // (groups of KT consecutive tiles, NB groups), and cleared
    constexpr int NT = G * IM * IN;
    constexpr int KT = (NT + 7) / 8, NB = NT / KT;
    static_assert(!BF || NB * KT == NT, "tile groups must cover the tiles");
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
            for (int jj = 0; jj < IN; ++jj)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[j][i][jj][v] = 0;   // 0 or 0.f

    // step state: q = index of step s within its k-block, sa = its A slot,
    // sb = the B slot of its window position s + gb - 1 (diagonal j = 0);
    // pa / pb[j] = fragment bases of the step, the zero block for diagonals
    // j < jmin (slice u = g0 + j - t < 0: the last G-1 steps of a k-block) and
    // for every diagonal of a padding step (q >= nA): no branch near the MFMAs
    int q = 0, sa = 0, sb = gb - 1;
    const int8_t *pa = fragA;
    const int8_t *pb[G];
    auto set_step = [&]() {
        const int t = t0 + q;
        const int jmin = (q < nA) ? max(0, t - g0) : G;
        pa = fragA + sa * BND_BLK;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            // slot (sb - j) mod RB, or RB (the zero block) when j < jmin;
            // integer masks instead of selects keep the step free of branches
            int sj = sb - j;
            sj += (sj >> 31) & RB;
            sj += (RB - sj) & -(int)(j < jmin);
            pb[j] = fragB + sj * BND_BLK;
        }
    };
    bnd_i4 a[2][IM], b[2][G][IN];
    // fragments of k-step ks; diagonals j < J0 are not read (no MFMA uses them)
    auto fetch = [&](int buf, int ks, auto jc) {
        constexpr int J0 = decltype(jc)::value;
#pragma unroll
        for (int i = 0; i < IMV; ++i) a[buf][i] = *(const bnd_i4 *)(pa + i * 4096 + koff[ks]);
#pragma unroll
        for (int j = J0; j < G; ++j)
#pragma unroll
            for (int jj = 0; jj < IN; ++jj) b[buf][j][jj] = *(const bnd_i4 *)(pb[j] + jj * 4096 + koff[ks]);
    };
    using J0All = std::integral_constant<int, 0>;
    // MFMAs of one k-step over the NT accumulator tiles (tile T = (j IM + i)
    // IN + jj) in the order O, O+1, ... mod NT, diagonals j >= J0 only
    auto mfma = [&](int buf, auto oc, auto jc) {
        constexpr int O = decltype(oc)::value, J0 = decltype(jc)::value;
#pragma unroll
        for (int idx = 0; idx < NT; ++idx) {
            const int T = (O + idx) % NT;
            const int j = T / (IM * IN), i = (T / IN) % IM, jj = T % IN;
            if (j < J0 || i >= IMV) continue;
            if constexpr (BF)
                acc[j][i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bnd_bf8, a[buf][i]), __builtin_bit_cast(bnd_bf8, b[buf][j][jj]), acc[j][i][jj],
                    0, 0, 0);
            else
                acc[j][i][jj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[buf][i], b[buf][j][jj], acc[j][i][jj], 0, 0, 0);
        }
    };
    using C0 = std::integral_constant<int, 0>;
    using None = std::integral_constant<int, -1>;

    // The barrier of step s sits inside its last k-step: wait for this
    // wave's loads of step s+1 (and for its own fragment reads), barrier,
    // refill the slots step s released with the loads of step s+1+PA, read
    // step s+1's first fragments, and only then run step s's last MFMAs,
    // which hide the latency of those reads.  The ring position p of a load
    // issued at barrier s is p = s+1+PA: slot p % RA = s % RA (step s's A
    // block); the B position s+PA+gb lands in slot s % RB (the first position
    // of step s's window, which step s+1 no longer reads).
    bnd_wait_vm<8 * (PA - 1)>();
    __builtin_amdgcn_s_barrier();
    issue_a();
    issue_b();
    set_step();
    fetch(0, 0, J0All{});
    // BF16 moves, a group of tiles at a time: a step with NX >= 0 runs group
    // NX first in its last k-step, adds its sums into tot and clears them
    // under the other tiles' MFMAs; in its first k-step the tiles of group
    // FT >= 0 (cleared by the step before) go last.  J0: the step's first
    // diagonal with a slice pair (the MFMAs of diagonals j < J0 would only
    // add zero products and are left out; J0 = G: a padding step, copies
    // and barrier only)
    auto step = [&](auto fc, auto nc, auto jc) {
        constexpr int FT = decltype(fc)::value, NX = decltype(nc)::value, J0 = decltype(jc)::value;
        // the scheduler places one fragment read (or one global->LDS copy)
        // after each MFMA, so no burst of reads or copies leaves the MFMA
        // pipe idle: NR = fragment reads, NM = MFMAs per k-step
        constexpr int NR = IMV + IN * (G - J0), NM = IMV * IN * (G - J0), NR0 = IMV + IN * G;
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            fetch((ks & 1) ^ 1, ks + 1, jc);
            if (ks == 0)
                mfma(0, std::integral_constant<int, ((FT + 1) * KT) % NT>{}, jc);
            else
                mfma(ks & 1, C0{}, jc);
#pragma unroll
            for (int r = 0; r < (NR < NM ? NR : NM); ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // LDS read
            }
            if constexpr (NM > NR) __builtin_amdgcn_sched_group_barrier(0x008, NM - NR, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        bnd_wait_vm<8 * (PA - 1)>();
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        if (++q == nsk) q = 0;
        if (++sa == RA) sa = 0;
        if (++sb == RB) sb = 0;
        set_step();
        fetch(0, 0, J0All{});
        issue_a();
        issue_b();
        mfma(1, std::integral_constant<int, (NX > 0 ? NX * KT : 0)>{}, jc);
        constexpr int NR1 = NR0 < NM ? NR0 : NM;
        constexpr int NC = (NM - NR1) < 8 ? (NM - NR1) : 8;
        if constexpr (BF && NX >= 0) {
#pragma unroll
            for (int e = 0; e < KT; ++e) {
                const int T = NX * KT + e;
                const int j = T / (IM * IN), i = (T / IN) % IM, jj = T % IN;
#pragma unroll
                for (int v = 0; v < 16; ++v) tot[j][i][jj][v] += (int)acc[j][i][jj][v];
                acc[j][i][jj] = AccT{};
            }
#pragma unroll
            for (int r = 0; r < NR1; ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // LDS read
            }
            // after the group's own MFMAs (KT <= NR): MFMA, copy and a share
            // of the 64 KT moves (read, convert, add; clear)
            constexpr int NV = (64 * KT + 7) / 8;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   // global->LDS copy
                __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);  // VALU
            }
        } else {
#pragma unroll
            for (int r = 0; r < NR1; ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // LDS read
            }
#pragma unroll
            for (int r = 0; r < NC; ++r) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   // global->LDS copy
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    if constexpr (BF) {
        // FP32 sums stay exact below 2^24, i.e. for 16 steps (1024 products
        // of at most 2^14).  Moving step R (0 <= R < NB) adds group R into
        // tot and clears it; the group gathers the MFMAs of the steps up to
        // its next move.  The steps run in segments of nm full steps -- the
        // moving steps R = 0, 1, .., NB-1 as often as they fit, the other
        // nm % NB without a move -- then the steps of diagonals J0 = jlo ..
        // jhi, then np padding steps.  An item with at least NB full steps
        // per k-block has a segment per k-block (nm = its full steps) and,
        // like the INT8 steps, leaves out the MFMAs of diagonals without a
        // slice pair: a group moved at its step r of a k-block gathers at
        // most NB-1-r moving, NB-1 other full and G-1 such steps there, and
        // r+1 steps of the next k-block, 2 NB + G - 2 in all.  Any other item
        // is one segment of nm = nsteps steps that run every MFMA (zero
        // blocks where no pair exists).
        static_assert(2 * NB + G - 2 <= 16, "groups must be moved within 16 steps");
        const int nfull = max(0, min(nA, g0 + 1 - t0));
        const bool skip = nfull >= NB;
        const int nseg = skip ? kb1 - kb0 : 1, nm = skip ? nfull : nsteps;
        const int jlo = skip ? max(1, t0 + nfull - g0) : G, jhi = skip ? min(G - 1, t1 - 1 - g0) : 0;
        const int np = skip ? nsk - nA : 0;
        for (int sg = 0; sg < nseg; ++sg) {
            int c = 0;
            for (; c + NB <= nm; c += NB)
                bnd_unroll(std::make_integer_sequence<int, NB>{}, [&](auto rc) {
                    constexpr int R = decltype(rc)::value;
                    step(std::integral_constant<int, (R + NB - 1) % NB>{}, rc, J0All{});
                });
            for (; c < nm; ++c) step(None{}, None{}, J0All{});
            bnd_unroll(std::make_integer_sequence<int, G - 1>{}, [&](auto rc) {
                constexpr int J = decltype(rc)::value + 1;
                if (J >= jlo && J <= jhi) step(None{}, None{}, std::integral_constant<int, J>{});
            });
            for (c = 0; c < np; ++c) step(None{}, None{}, std::integral_constant<int, G>{});
        }
    } else {
        // per k-block: nfull steps with a slice pair on every diagonal, the
        // steps whose first J0 = jlo .. jhi diagonals have none (slice t of
        // A above g0), and npad padding steps, which only refill the B window
        const int nfull = max(0, min(nA, g0 + 1 - t0));
        const int jlo = max(1, t0 + nfull - g0), jhi = min(G - 1, t1 - 1 - g0);
        const int npad = nsk - nA;
        // (a second instantiation of these steps for a last, partial
        // k-block with fewer k-steps made the compiler spill the
        // accumulators: every k-block runs all four k-steps)
        for (int kb = kb0; kb < kb1; ++kb) {
            for (int c = 0; c < nfull; ++c) step(None{}, None{}, J0All{});
            bnd_unroll(std::make_integer_sequence<int, G - 1>{}, [&](auto rc) {
                constexpr int J = decltype(rc)::value + 1;
                if (J >= jlo && J <= jhi) step(None{}, None{}, std::integral_constant<int, J>{});
            });
            for (int c = 0; c < npad; ++c) step(None{}, None{}, std::integral_constant<int, G>{});
        }
    }
    // no global->LDS copy may still be in flight when the workgroup ends
    bnd_wait_vm<0>();

    // acc element v of a 32 x 32 tile: row 8 (v / 4) + 4 (lane / 32) + v % 4,
    // column lane % 32; the band's G sums of an entry go out as one INT64,
    // sum_j D_j 2^(8 (G-1-j)) (a diagonal this chunk has no pair for, or one
    // below 0, holds 0)
    const int nbands = (S + G - 1) / G;
    const int band = (g0 - (S - G * nbands)) / G;
    long long *plane =
        D + (long long)(bnd_first_band_plane(S, G, nkb, cap, band) + kb0 / bnd_band_kblocks(S, G, cap, band)) *
                ((long long)m * n);
    // the lane index again from the exec mask, so that no lane-dependent
    // value has to live (in scratch memory) across the main loop
    const int el = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
#pragma unroll
    for (int i = 0; i < IMV; ++i)
#pragma unroll
        for (int jj = 0; jj < IN; ++jj) {
            const long long col = col0 + wn * TNW + jj * 32 + (el & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long row = row0 + wm * TMW + i * 32 + (v >> 2) * 8 + (el >> 5) * 4 + (v & 3);
                if (row < m && col < n) {
                    long long P = 0;
#pragma unroll
                    for (int j = 0; j < G; ++j) {
                        int dj;
                        if constexpr (BF)
                            dj = tot[j][i][jj][v] + (int)acc[j][i][jj][v];
                        else
                            dj = acc[j][i][jj][v];
                        P = P * 256 + dj;
                    }
                    plane[row * n + col] = P;
                }
            }
        }
}


// Single problem.  Grid: x = output tiles (tiles_m * tiles_n), y = band items.
template <int G, int PA, int WM>
__global__ __launch_bounds__(256) void band_i8_kernel(const int8_t *__restrict__ Ab, const int8_t *__restrict__ Bb,
                                                      int S, int nkb, int cap, int tiles_m, int tiles_n, int m, int n,
                                                      BandTable tab, long long *__restrict__ D)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[bnd_lds_bytes<G, PA>()];
    const BandItem it = tab.item[blockIdx.y];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const long long tstride = (long long)nkb * S * BND_BLK;
    band_run<G, PA, WM>(lds, Ab + tm * tstride, Bb + tn * tstride, S, nkb, cap, it.g0, 0, it.g0 + G, it.kb0,
                        it.kb1, (long long)tm * BND_TM, (long long)tn * BND_TM, m, n, D);
}

// The same over BF16 slices (64 bf16 per block row, nkb = kp / 64).
template <int G, int PA, int WM>
__global__ __launch_bounds__(256) void band_bf16_kernel(const int8_t *__restrict__ Ab, const int8_t *__restrict__ Bb,
                                                        int S, int nkb, int cap, int tiles_m, int tiles_n, int m, int n,
                                                        BandTable tab, long long *__restrict__ D)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[bnd_lds_bytes<G, PA>()];
    const BandItem it = tab.item[blockIdx.y];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const long long tstride = (long long)nkb * S * BND_BLK;
    band_run<G, PA, WM, true>(lds, Ab + tm * tstride, Bb + tn * tstride, S, nkb, cap, it.g0, 0, it.g0 + G,
                              it.kb0, it.kb1, (long long)tm * BND_TM, (long long)tn * BND_TM, m, n, D);
}

// Band items: bands of G diagonals, each band's k-blocks cut into K-chunks
// (bnd_band_kblocks; work = steps, one per slice of A and k-block).  When S
// is not a multiple of G, the S % G leftover diagonals are the cheapest ones,
// 0 .. S%G - 1 (one slice pair for diagonal 0: a band that reuses its
// operands less should carry little work), and they run as a G-wide band
// shifted down, g0 = S%G - G < 0: its first G - S%G diagonals are empty.
// band_run handles that band unchanged -- slice u = g0 + j - t < 0 reads the
// zero block, the INT8 step types leave those MFMAs out, the epilogue skips
// diagonals below 0 -- so every item is G wide and one launch of the G
// kernel covers all of them.
struct BandWork {
    BandItem it;
    long long work;
};

inline void band_items(int S, int G, int nkb, int cap, std::vector<BandWork> &items)
{
    items.clear();
    const int nb = (S + G - 1) / G;
    for (int q = 0; q < nb; ++q) {
        const int g0 = S - G * (nb - q);
        const int kc = bnd_band_kblocks(S, G, cap, q);
        for (int kb0 = 0; kb0 < nkb; kb0 += kc) {
            const int kb1 = std::min(kb0 + kc, nkb);
            items.push_back(BandWork{BandItem{(short)g0, (short)G, (short)kb0, (short)kb1},
                                     (long long)(kb1 - kb0) * (g0 + G)});
        }
    }
}

}  // namespace mdg
