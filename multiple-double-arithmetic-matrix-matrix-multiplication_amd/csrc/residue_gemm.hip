// residue_gemm.hip -- multi-double GEMM from INT8 GEMMs of residues
// (Ozaki, Uchino, Imamura, "Ozaki Scheme II: a GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique",
// 2025), extended to the hundreds of bits of quad- to hexa-double.
//
// Like the slice GEMMs, every row of A / column of B gets a scale exponent
// and every entry becomes the fixed-point integer a = floor(x 2^(W-2-E))
// (slices.hpp).  The integer product Z = sum_k a_rk b_kc is then computed
// exactly from its residues modulo pairwise coprime moduli m_1..m_n with
// M = prod m_i > 4 max|Z|:
//   * narrow moduli (m <= 256): the balanced residues of a and b are INT8,
//     and one INT8 x INT8 -> INT32 GEMM gives Z mod m;
//   * wide moduli (odd primes < 64768): the balanced residue is cut into two
//     INT8 digits, r = 256 hi + lo, and three GEMMs (hi*hi, lo*lo, and
//     hi*lo + lo*hi as one GEMM of depth 2k) give Z mod m.
// The slice GEMMs need S(S+1)/2 slice-pair products for S ~ W/8 slices;
// this needs about (2W + log2 k)/8 narrow and 4 GEMM units per 16 further
// bits -- at quad-double 81 INT8 GEMMs of depth k instead of 435.
//
// Reconstruction (per output entry, exact integer arithmetic):
//   Z = sum_i r_i c_i - q M,  c_i = (M/m_i) * ((M/m_i)^-1 mod m_i),
//   q = round(sum_i r_i y_i / m_i)  (in FP64; |Z| <= M/4 makes it exact),
// accumulated in 32-bit words with 64-bit accumulators, carried, and folded
// into L limbs from the most significant word with the TwoSum inc_d cascade
// (C++/TCcuda.cu:876-906 folds its parts the same way).
//
// Error: the only roundings are the truncation of each entry to W bits and
// the final fold, so the normwise error is about L 2^(3-W+spread_a+spread_b);
// W = target + spread_a + spread_b + 3 + log2(L).
#include <climits>
#include <cmath>
#include <cstring>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "i8tile.hpp"
#include "md_arith.hpp"
#include "recon.hpp"
#include "slices.hpp"
#include "split_common.hpp"

namespace mdg {

constexpr int RES_NC = 32;       // 37-bit chunks of the widest fixed-point integer (36 words)
constexpr int RES_NC24 = 48;     // its 24-bit chunks
constexpr int RES_KC = 130944;   // GEMM depth with exact INT32 sums: RES_KC * 2^14 < 2^31, multiple of 128
constexpr int RES_TK = 64;       // B residue tile depth (k rows)

// Device view of one modulus set (n = n8 + n16 moduli, narrow ones first).
// All tables hold exact integers as doubles: the kernels evaluate them in
// FP64 FMA chains whose partial sums stay below 2^53, hence exact.
struct ResTab {
    int n8, n16;                // narrow / wide moduli
    int ng8;                    // narrow groups: triples of narrow moduli (the last may be shorter)
    int nw;                     // 24-bit words per CRT constant (those of M; Z keeps its sign in the carry)
    const double *md;           // m_i                                                  [n]
    const double *inv;          // 1 / m_i                                              [n]
    const double *pw37;         // balanced 2^(37c) mod m_i                              [n8][RES_NC]
    const double *pw37h;        // wide m_j: balanced 2^(37c) mod m_j = 256 hi + lo, hi   [n16][RES_NC]
    const double *pw37l;        //                                                lo     [n16][RES_NC]
    const double *gP;           // group moduli: narrow triple products P_g, then the wide moduli   [G]
    const double *ginv;         // 1 / P_g                                              [G]
    const double *gfrac;        // y_g / P_g, y_g = (M/P_g)^-1 mod P_g                  [G]
    const double *gE;           // narrow group g: E_x = (P/m_x)((P/m_x)^-1 mod m_x) of its members, 0 if absent [ng8][3]
    const double *cw;           // c_g = (M/P_g) y_g, [G][nw] 24-bit words, least significant first
    const double *Mw;           // M, [nw] 24-bit words
    // FP64-MFMA moduli (odd primes below 2^21; then n8 = n16 = 0, G = nf)
    int nf;
    const double *pw24;         // balanced 2^(24c) mod m_i                              [nf][RES_NC24]
    const double *c32;          // balanced 2^32 mod m_i                                 [nf]
};

// v mod m in [0, m) for an integer-valued double |v| < 2^53: the FP64
// quotient is off by at most one and the remainder fma is exact
__device__ __forceinline__ double mod_pos(double v, double md, double inv)
{
    const double q = floor(v * inv);
    double r = fma(-q, md, v);
    r += (r < 0.0) ? md : 0.0;
    r -= (r >= md) ? md : 0.0;
    return r;
}

// bits [37c, 37c + 37) of the non-negative NW-word integer N, as a double
// (32-bit funnel shifts only: a 64-bit combination of two array words
// would be turned into a 64-bit load and keep N in scratch memory)
template <int NW>
__device__ __forceinline__ double chunk37(const uint32_t (&N)[NW], int c)
{
    const int pos = 37 * c, j = pos >> 5, off = pos & 31;
    const uint32_t w0 = N[j], w1 = (j + 1 < NW) ? N[j + 1] : 0u, w2 = (j + 2 < NW) ? N[j + 2] : 0u;
    const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, off);
    const uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, off) & 31u;
    return fma((double)hi, 4294967296.0, (double)lo);
}

template <int EPT> struct PackOf;
template <> struct PackOf<4> { typedef uint32_t T; };
template <> struct PackOf<2> { typedef uint16_t T; };
template <> struct PackOf<1> { typedef uint8_t T; };

// Balanced residue of an integer-valued double v, |v| < 2^51, modulo an odd
// m: fl(v * fl(1/m)) is within |v|/m 2^-52 < 1/(2m) of v/m, and v/m is at
// least 1/(2m) from every half-integer (m odd), so rint gives the nearest
// quotient and the remainder fma is exact: r in [-(m-1)/2, (m-1)/2].
__device__ __forceinline__ double mod_bal_odd(double v, double md, double inv)
{
    return fma(-rint(v * inv), md, v);
}

// |floor(x 2^shift)| as NC 37-bit chunks; returns the sign and the low byte
// of the two's-complement value (its residue modulo 256)
template <int L, int NWU, int NC>
__device__ __forceinline__ double prep_entry(const double (&x)[L], int shift, double (&ch)[NC], uint32_t &low)
{
    uint32_t N[NWU];
    to_fixed<L, NWU>(x, shift, N);
    low = N[0] & 255u;
    const bool neg = (int32_t)N[NWU - 1] < 0;
    if (neg) {
        uint32_t cy = 1;
#pragma unroll
        for (int j = 0; j < NWU; ++j) {
            const uint32_t v = ~N[j] + cy;
            cy = (cy && v == 0u) ? 1u : 0u;
            N[j] = v;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) ch[c] = chunk37<NWU>(N, c);
    return neg ? -1.0 : 1.0;
}

// Balanced residues of EPT entries that share one scale shift, packed EPT
// bytes per plane:
//   narrow plane i at o8 + i * plane8                       r in [-128, 127]
//   wide modulus j at o16 + j * plane16: hi at +0, lo at +kp (and hi again
//   at +2kp when triple), r = 256 hi + lo with hi, lo in [-128, 127].
// |N| = |floor(x 2^shift)| is cut into at most 32 chunks of 37 bits; with
// |2^(37c) mod m| (balanced) below 2^7 -- for a wide modulus split as
// 256 hi + lo into two such tables -- every partial sum stays below 2^47,
// exact in FP64 and inside mod_bal_odd's range.  Modulus 256 is the low
// byte of N itself.
template <int L, int NWU, int EPT>
__device__ __forceinline__ void entry_residues(const double (&x)[EPT][L], int shift, const ResTab &t, int8_t *o8,
                                               long long plane8, int8_t *o16, long long plane16, int kp, bool triple)
{
    constexpr int NC = (32 * NWU + 36) / 37;
    typedef typename PackOf<EPT>::T PT;
    double ch[EPT][NC];
    double sg[EPT];
    uint32_t low[EPT];
    // entries prepared one by one in straight-line code (a loop over them
    // would not be unrolled and would put the arrays in scratch memory)
    sg[0] = prep_entry<L, NWU, NC>(x[0], shift, ch[0], low[0]);
    if constexpr (EPT > 1) sg[1] = prep_entry<L, NWU, NC>(x[1], shift, ch[1], low[1]);
    if constexpr (EPT > 2) {
        sg[2] = prep_entry<L, NWU, NC>(x[2], shift, ch[2], low[2]);
        sg[3] = prep_entry<L, NWU, NC>(x[3], shift, ch[3], low[3]);
    }
    int i = 0;
    if (t.n8 > 0 && t.md[0] == 256.0) {
        PT word = 0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) word |= (PT)(low[e] << (8 * e));
        *(PT *)o8 = word;
        i = 1;
    }
    // independent FMA chains per entry: with fewer entries per thread each
    // chain is split into partial sums (exact integers, so the same v)
    constexpr int CH = EPT >= 4 ? 1 : (EPT == 2 ? 2 : 4);
    constexpr int CHW = EPT >= 2 ? 1 : 2;
    for (; i < t.n8; ++i) {
        const double *pw = t.pw37 + i * RES_NC;
        const double md = t.md[i], inv = t.inv[i];
        PT word = 0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            double vp[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) vp[q] = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) vp[c % CH] = fma(ch[e][c], pw[c], vp[c % CH]);
            double v = vp[0];
#pragma unroll
            for (int q = 1; q < CH; ++q) v += vp[q];
            const int r = (int)mod_bal_odd(v * sg[e], md, inv);
            word |= (PT)(((uint32_t)r & 255u) << (8 * e));
        }
        *(PT *)(o8 + i * plane8) = word;
    }
    for (int j = 0; j < t.n16; ++j) {
        const double *ph = t.pw37h + j * RES_NC;
        const double *pl = t.pw37l + j * RES_NC;
        const double md = t.md[t.n8 + j], inv = t.inv[t.n8 + j];
        PT hw = 0, lw = 0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            double vhp[CHW], vlp[CHW];
#pragma unroll
            for (int q = 0; q < CHW; ++q) vhp[q] = vlp[q] = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                vhp[c % CHW] = fma(ch[e][c], ph[c], vhp[c % CHW]);
                vlp[c % CHW] = fma(ch[e][c], pl[c], vlp[c % CHW]);
            }
            double vh = vhp[0], vl = vlp[0];
#pragma unroll
            for (int q = 1; q < CHW; ++q) {
                vh += vhp[q];
                vl += vlp[q];
            }
            const double r1 = mod_bal_odd(vh * sg[e], md, inv);                  // |r1| < 2^15
            const int r = (int)mod_bal_odd(fma(r1, 256.0, vl * sg[e]), md, inv);  // < 2^47 in
            const int lo = ((r + 128) & 255) - 128;
            const int hi = (r - lo) >> 8;
            hw |= (PT)(((uint32_t)hi & 255u) << (8 * e));
            lw |= (PT)(((uint32_t)lo & 255u) << (8 * e));
        }
        int8_t *o = o16 + j * plane16;
        *(PT *)o = hw;
        *(PT *)(o + kp) = lw;
        if (triple) *(PT *)(o + 2 * kp) = hw;
    }
}

// the residues of EPT zero entries (padding, zero rows and columns): zero
// words in every plane, in the layout of entry_residues
template <int EPT>
__device__ __forceinline__ void zero_residues(const ResTab &t, int8_t *o8, long long plane8, int8_t *o16,
                                              long long plane16, int kp, bool triple)
{
    typedef typename PackOf<EPT>::T PT;
    for (int i = 0; i < t.n8; ++i) *(PT *)(o8 + i * plane8) = 0;
    for (int j = 0; j < t.n16; ++j) {
        int8_t *o = o16 + j * plane16;
        *(PT *)o = 0;
        *(PT *)(o + kp) = 0;
        if (triple) *(PT *)(o + 2 * kp) = 0;
    }
}

// entries per thread: packed stores of 4/2/1 bytes per plane, within the
// register budget of the fixed-point words and chunks
template <int L>
struct ResEPT {
    static constexpr int value = (L <= 4) ? 4 : (L == 8) ? 2 : 1;
};

// A: thread per EPT consecutive entries (r, kk..kk+EPT-1) of the padded
// mp x kp domain.  Narrow planes R8[i][r][kk] (mp x kp), wide planes
// R16[j][r][2kp]: hi at kk, lo at kp + kk.
template <int L, int NWU>
__device__ __forceinline__ void residue_a_at(long long idx, int m, int k, int mp, int kp, const double *__restrict__ A,
                                             long long lda, const int *__restrict__ E, int W, const ResTab &t,
                                             int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    constexpr int EPT = ResEPT<L>::value;
    const int kq = kp / EPT;
    if (idx >= (long long)mp * kq) return;
    const long long r = idx / kq;
    const int kk = (int)(idx - r * kq) * EPT;
    double x[EPT][L];
    int shift = 0;
    const bool live = r < m && E[r] != ZERO_SCALE;
    if (!live || kk >= k) {
        zero_residues<EPT>(t, R8 + r * kp + kk, (long long)mp * kp, R16 + r * 2 * kp + kk, 2ll * mp * kp, kp, false);
        return;
    }
    shift = W - 2 - E[r];
#pragma unroll
    for (int e = 0; e < EPT; ++e)
#pragma unroll
        for (int l = 0; l < L; ++l) x[e][l] = (kk + e < k) ? A[(r * lda + kk + e) * L + l] : 0.0;
    entry_residues<L, NWU, EPT>(x, shift, t, R8 + r * kp + kk, (long long)mp * kp, R16 + r * 2 * kp + kk,
                                2ll * mp * kp, kp, false);
}

template <int L, int NWU>
__global__ __launch_bounds__(256) void residue_a_kernel(int m, int k, int mp, int kp, const double *__restrict__ A,
                                                        long long lda, const int *__restrict__ E, int W, ResTab t,
                                                        int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    residue_a_at<L, NWU>((long long)blockIdx.x * blockDim.x + threadIdx.x, m, k, mp, kp, A, lda, E, W, t, R8, R16);
}

// B: a block stages a RES_TK x TC tile of B (rows kk, columns c; coalesced
// along c) in LDS, then each thread takes EPT consecutive kk of one column,
// so the transposed residue writes coalesce along kk.  Narrow planes
// R8[i][c][kk] (np x kp), wide planes R16[j][c][3kp]: hi at kk and
// 2kp + kk, lo at kp + kk -- [kp, 3kp) is (lo, hi) against A's (hi, lo).
// blk numbers the (np / TC) x (kp / RES_TK) tiles.
template <int L, int NWU>
__device__ __forceinline__ void residue_b_at(long long blk, int k, int n, int kp, int np, const double *__restrict__ B,
                                             long long ldb, const int *__restrict__ F, int W, const ResTab &t,
                                             int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    constexpr int EPT = ResEPT<L>::value;
    constexpr int TPC = RES_TK / EPT;   // threads per column
    constexpr int TC = 256 / TPC;       // columns per tile
    constexpr int ROW = TC * L + 1;     // padded LDS row, doubles
    __shared__ double tile[RES_TK * ROW];
    const int ktiles = kp / RES_TK;
    const int c0 = (int)(blk / ktiles) * TC;
    const int k0 = (int)(blk % ktiles) * RES_TK;
    for (int q = threadIdx.x; q < RES_TK * TC * L; q += 256) {
        const int row = q / (TC * L), col = q - row * (TC * L);
        const int kk = k0 + row, c = c0 + col / L;
        tile[row * ROW + col] = (kk < k && c < n) ? B[((long long)kk * ldb + c0) * L + col] : 0.0;
    }
    __syncthreads();
    const int cc = threadIdx.x / TPC, kq = threadIdx.x % TPC;
    const int c = c0 + cc, kk = k0 + kq * EPT;
    const bool live = c < n && F[c] != ZERO_SCALE;
    if (!live || kk >= k) {
        zero_residues<EPT>(t, R8 + (long long)c * kp + kk, (long long)np * kp, R16 + (long long)c * 3 * kp + kk,
                           3ll * np * kp, kp, true);
        return;
    }
    const int shift = W - 2 - F[c];
    double x[EPT][L];
#pragma unroll
    for (int e = 0; e < EPT; ++e)
#pragma unroll
        for (int l = 0; l < L; ++l) x[e][l] = tile[(kq * EPT + e) * ROW + cc * L + l];
    entry_residues<L, NWU, EPT>(x, shift, t, R8 + (long long)c * kp + kk, (long long)np * kp,
                                R16 + (long long)c * 3 * kp + kk, 3ll * np * kp, kp, true);
}

template <int L, int NWU>
__global__ __launch_bounds__(256) void residue_b_kernel(int k, int n, int kp, int np, const double *__restrict__ B,
                                                        long long ldb, const int *__restrict__ F, int W, ResTab t,
                                                        int8_t *__restrict__ R8, int8_t *__restrict__ R16)
{
    residue_b_at<L, NWU>(blockIdx.x, k, n, kp, np, B, ldb, F, W, t, R8, R16);
}

// ----------------------------------------------- FP64-MFMA residue planes --
// bits [24c, 24c + 24) of the non-negative NW-word integer N, as a double
template <int NW>
__device__ __forceinline__ double chunk24(const uint32_t (&N)[NW], int c)
{
    const int pos = 24 * c, j = pos >> 5, off = pos & 31;
    const uint32_t w0 = N[j], w1 = (j + 1 < NW) ? N[j + 1] : 0u;
    return (double)(__builtin_amdgcn_alignbit(w1, w0, off) & 0xFFFFFFu);
}

// Balanced residues of one entry modulo the FP64 moduli (|r| < 2^20), one
// INT32 plane per modulus at o + i * plane: 24-bit chunks of |N| against
// balanced 2^(24c) mod m (< 2^20): at most 48 terms of 2^44, below 2^50.
template <int L, int NWU>
__device__ __forceinline__ void f64_entry_residues(const double (&x)[L], int shift, const ResTab &t, int32_t *o,
                                                   long long plane)
{
    constexpr int NC = (32 * NWU + 23) / 24;
    uint32_t N[NWU];
    to_fixed<L, NWU>(x, shift, N);
    const bool neg = (int32_t)N[NWU - 1] < 0;
    if (neg) {
        uint32_t cy = 1;
#pragma unroll
        for (int j = 0; j < NWU; ++j) {
            const uint32_t v = ~N[j] + cy;
            cy = (cy && v == 0u) ? 1u : 0u;
            N[j] = v;
        }
    }
    const double sg = neg ? -1.0 : 1.0;
    double ch[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) ch[c] = sg * chunk24<NWU>(N, c);
    for (int i = 0; i < t.nf; ++i) {
        const double *pw = t.pw24 + i * RES_NC24;
        // four independent chains (exact integer partial sums)
        double vp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < NC; ++c) vp[c & 3] = fma(ch[c], pw[c], vp[c & 3]);
        const double v = (vp[0] + vp[1]) + (vp[2] + vp[3]);
        o[i * plane] = (int32_t)mod_bal_odd(v, t.md[i], t.inv[i]);
    }
}

// A: thread per entry (r, kk) of the padded mp x kp domain; planes
// R[i][r][kk] (mp x kp)
template <int L, int NWU>
__global__ __launch_bounds__(256) void f64_residue_a_kernel(int m, int k, int mp, int kp, const double *__restrict__ A,
                                                            long long lda, const int *__restrict__ E, int W, ResTab t,
                                                            int32_t *__restrict__ R)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)mp * kp) return;
    const long long r = idx / kp;
    const int kk = (int)(idx - r * kp);
    const bool live = r < m && kk < k && E[r] != ZERO_SCALE;
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = live ? A[(r * lda + kk) * L + l] : 0.0;
    f64_entry_residues<L, NWU>(x, live ? W - 2 - E[r] : 0, t, R + r * kp + kk, (long long)mp * kp);
}

// B: a block stages a RES_TK x 4 tile of B in LDS (coalesced along the
// columns), then each wave takes one column and 64 consecutive kk, so the
// transposed planes R[i][c][kk] (np x kp) are written 256 bytes at a time.
template <int L, int NWU>
__global__ __launch_bounds__(256) void f64_residue_b_kernel(int k, int n, int kp, int np, const double *__restrict__ B,
                                                            long long ldb, const int *__restrict__ F, int W, ResTab t,
                                                            int32_t *__restrict__ R)
{
    constexpr int TC = 4, ROW = TC * L + 1;
    __shared__ double tile[RES_TK * ROW];
    const int ktiles = kp / RES_TK;
    const int c0 = (int)(blockIdx.x / ktiles) * TC;
    const int k0 = (int)(blockIdx.x % ktiles) * RES_TK;
    for (int q = threadIdx.x; q < RES_TK * TC * L; q += 256) {
        const int row = q / (TC * L), col = q - row * (TC * L);
        const int kk = k0 + row, c = c0 + col / L;
        tile[row * ROW + col] = (kk < k && c < n) ? B[((long long)kk * ldb + c0) * L + col] : 0.0;
    }
    __syncthreads();
    const int cc = threadIdx.x >> 6, kq = threadIdx.x & 63;
    const int c = c0 + cc, kk = k0 + kq;
    const bool live = c < n && F[c] != ZERO_SCALE;
    double x[L];
#pragma unroll
    for (int l = 0; l < L; ++l) x[l] = live ? tile[kq * ROW + cc * L + l] : 0.0;
    f64_entry_residues<L, NWU>(x, live ? W - 2 - F[c] : 0, t, R + (long long)c * kp + kk, (long long)np * kp);
}

// C[r][c] from the INT32 planes: narrow D8[i][q] (chunk q of nck1), wide
// Dhh[j][q], Dll[j][q] (nck1 chunks) and Dx[j][q] (nck2 chunks); plane p of a
// block at p * ps.  Two-level CRT in exact FP64:
//   * narrow triple g: R_g = (v_a E_a + v_b E_b + v_c E_c) mod P_g, P_g < 2^24
//     (the plane values v enter directly while |v| E stays below 2^51,
//     i.e. depth <= 8192 in one chunk; otherwise each v is reduced first);
//   * wide modulus j: R = (hh 2^16 + xx 2^8 + ll) mod m_j;
//   * Z = sum_g R_g c_g - q M in 24-bit word columns (R_g c_gw < 2^48, at
//     most 17 narrow groups and 192 wide moduli: sums < 2^52.2, exact).
// Z is folded into L limbs from its leading 48-bit piece: the first L + 1
// pieces through fast_renorm, the rest through inc_d.
// narrow groups whose plane loads are issued together: one -- the registers
// saved buy occupancy, which hides the load latency better (qd n=4096:
// 1.42 ms against 1.68 ms with four and 1.90 ms with eight)
constexpr int RES_GB = 1;
// octo- and hexa-double: the CRT words take most registers (occupancy 1-2
// waves), so one group of loads at a time leaves every plane load's latency
// exposed; four groups go out together (hd n=4096: 25.1 ms against 28.9 ms
// with one and 49.9 ms with eight).  The FP64-modulus planes keep pairs.
constexpr int RES_GB_WIDE = 4;
template <int L>
constexpr int res_gb()
{
    return L <= 4 ? RES_GB : RES_GB_WIDE;
}

template <int NW>
__device__ __forceinline__ void crt_acc(double (&acc)[NW], double &tq, const ResTab &t, int g, double R)
{
    tq = fma(R, t.gfrac[g], tq);
    const double *cw = t.cw + (long long)g * NW;
#pragma unroll
    for (int w = 0; w < NW; ++w) acc[w] = fma(R, cw[w], acc[w]);
}

// the GEMM result planes of one call
struct ResPlanes {
    const int *D8, *Dhh, *Dll, *Dx;   // INT8 moduli: INT32 planes
    const long long *D64;             // FP64 moduli: INT64 planes
    long long ps;                     // plane stride (elements)
    int nck1, nck2;                   // K-chunks of the depth-k and depth-2k GEMMs
    int reduce;                       // narrow triples: reduce each plane value first
    // planes already reduced to [0, m) (the fused batch's GEMM epilogue):
    // narrow i at P8 + i ps; wide j: hh at P16 + j ps, ll at (n16 + j) ps,
    // cross at (2 n16 + j) ps; one K-chunk
    const uint8_t *P8 = nullptr;
    const uint16_t *P16 = nullptr;
};

template <int L, int NW, bool F64, bool PRE = false>
__device__ __forceinline__ void residue_reconstruct_at(long long idx, int m, int n, int W, const ResTab &t,
                                                       const ResPlanes &pl, const int *__restrict__ E,
                                                       const int *__restrict__ F, double *__restrict__ C,
                                                       long long ldc)
{
    const int *__restrict__ D8 = pl.D8;
    const int *__restrict__ Dhh = pl.Dhh;
    const int *__restrict__ Dll = pl.Dll;
    const int *__restrict__ Dx = pl.Dx;
    const long long *__restrict__ D64 = pl.D64;
    const long long ps = pl.ps;
    const int nck1 = pl.nck1, nck2 = pl.nck2, reduce = pl.reduce;
    constexpr int GB = res_gb<L>();
    if (idx >= (long long)m * n) return;
    const long long r = idx / n, c = idx - r * n;
    double y[L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    const int er = E[r], fc = F[c];
    if (er != ZERO_SCALE && fc != ZERO_SCALE) {
        double acc[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) acc[w] = 0.0;
        double tq = 0.0;
        if constexpr (F64) {
            // INT64 planes of the FP64 moduli, |v| <= k 2^40: v = vh 2^32 + vl,
            // v mod m = ((vh mod m) (2^32 mod m) + vl) mod m, all exact in FP64
            for (int i0 = 0; i0 < t.nf; i0 += 2 * RES_GB) {
                long long v[2 * RES_GB];
#pragma unroll
                for (int u = 0; u < 2 * RES_GB; ++u) {
                    const long long val = D64[(long long)min(i0 + u, t.nf - 1) * ps + idx];
                    v[u] = (i0 + u < t.nf) ? val : 0;
                }
#pragma unroll
                for (int u = 0; u < 2 * RES_GB; ++u) {
                    const int i = i0 + u;
                    if (i >= t.nf) break;
                    const double r1 = mod_bal_odd((double)(v[u] >> 32), t.md[i], t.inv[i]);
                    const double lo = (double)(uint32_t)(v[u] & 0xFFFFFFFFll);
                    crt_acc<NW>(acc, tq, t, i, mod_pos(fma(r1, t.c32[i], lo), t.md[i], t.inv[i]));
                }
            }
        }
        for (int g0 = 0; g0 < t.ng8; g0 += GB) {
            int v[GB][3];
#pragma unroll
            for (int u = 0; u < GB; ++u)
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    // unconditional load of a clamped plane, then a select:
                    // a branch around each load would make each one wait
                    const int i = 3 * (g0 + u) + x, ic = min(i, t.n8 - 1);
                    int val;
                    if constexpr (PRE)
                        val = (int)pl.P8[(long long)ic * ps + idx];
                    else
                        val = D8[(long long)ic * nck1 * ps + idx];
                    v[u][x] = (i < t.n8) ? val : 0;
                }
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const int g = g0 + u;
                if (g >= t.ng8) break;
                const double *Eg = t.gE + 3 * g;
                double V;
                if (!reduce) {
                    V = fma((double)v[u][0], Eg[0], fma((double)v[u][1], Eg[1], (double)v[u][2] * Eg[2]));
                } else {
                    V = 0.0;
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const int i = 3 * g + x;
                        if (i >= t.n8) break;
                        long long sum = v[u][x];
                        for (int q = 1; q < nck1; ++q) sum += D8[((long long)i * nck1 + q) * ps + idx];
                        V = fma(mod_pos((double)sum, t.md[i], t.inv[i]), Eg[x], V);
                    }
                }
                crt_acc<NW>(acc, tq, t, g, mod_pos(V, t.gP[g], t.ginv[g]));
            }
        }
        for (int j0 = 0; j0 < t.n16; j0 += GB) {
            int hh[GB], ll[GB], xx[GB];
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const bool on = j0 + u < t.n16;
                const long long j = min(j0 + u, t.n16 - 1);
                int h0, l0, x0;
                if constexpr (PRE) {
                    h0 = (int)pl.P16[j * ps + idx];
                    l0 = (int)pl.P16[(t.n16 + j) * ps + idx];
                    x0 = (int)pl.P16[(2 * t.n16 + j) * ps + idx];
                } else {
                    h0 = Dhh[j * nck1 * ps + idx];
                    l0 = Dll[j * nck1 * ps + idx];
                    x0 = Dx[j * nck2 * ps + idx];
                }
                hh[u] = on ? h0 : 0;
                ll[u] = on ? l0 : 0;
                xx[u] = on ? x0 : 0;
            }
#pragma unroll
            for (int u = 0; u < GB; ++u) {
                const int j = j0 + u;
                if (j >= t.n16) break;
                long long h = hh[u], l = ll[u], x = xx[u];
                for (int q = 1; q < nck1; ++q) {
                    h += Dhh[((long long)j * nck1 + q) * ps + idx];
                    l += Dll[((long long)j * nck1 + q) * ps + idx];
                }
                for (int q = 1; q < nck2; ++q) x += Dx[((long long)j * nck2 + q) * ps + idx];
                const int g = t.ng8 + j;   // h 2^16 + x 2^8 + l: exact, < 2^53
                crt_acc<NW>(acc, tq, t, g,
                            mod_pos(fma((double)h, 65536.0, fma((double)x, 256.0, (double)l)), t.gP[g], t.ginv[g]));
            }
        }
        // Z = sum R_g c_g - q M in two's-complement 24-bit words.  The carry
        // chain stays in FP64: every v below is an integer with |v| <
        // 2^52.2 + 2^29 < 2^53 (the column sums' bound plus a carry), so
        // floor(v 2^-24) and v - 2^24 floor(v 2^-24) are exact -- the words
        // and carries of an INT64 chain, without its conversions
        const double q = rint(tq);
        double z[NW];
        double carry = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const double v = fma(-q, t.Mw[w], acc[w]) + carry;
            carry = floor(v * 0x1p-24);
            z[w] = fma(-carry, 16777216.0, v);
        }
        const bool neg = carry < 0.0;
        if (neg) {
            // |Z| = the two's complement: (2^24 - 1 - z) words plus one
            double cy = 1.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const double v = (16777215.0 - z[w]) + cy;
                cy = (v >= 16777216.0) ? 1.0 : 0.0;
                z[w] = v - 16777216.0 * cy;
            }
        }
        // 48-bit pieces, the leading one first: K pieces carry at least
        // 1 + 48 (K - 1) >= 53 L + 20 bits of |Z|
        constexpr int NP = NW / 2;
        constexpr int K = (53 * L + 19) / 48 + 2;
        double d[NP];
        int tp = -1;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            d[p] = fma(z[2 * p + 1], 16777216.0, z[2 * p]);
            tp = (d[p] != 0.0) ? p : tp;
        }
        if (tp >= 0) {
            const int sc = er + fc + 4 - 2 * W;
            double piece[K];
            top_pieces<NP, K>(d, tp, piece);
#pragma unroll
            for (int s = 0; s < K; ++s) piece[s] = ldexp(piece[s], 48 * (tp - s) + sc);
            fast_renorm<L>(piece, y);
#pragma unroll
            for (int s = L + 1; s < K; ++s) md_inc_d<L>(y, piece[s]);
            if (neg) {
#pragma unroll
                for (int l = 0; l < L; ++l) y[l] = -y[l];
            }
        }
    }
    double *o = C + (r * ldc + c) * L;
#pragma unroll
    for (int l = 0; l < L; ++l) o[l] = y[l];
}

template <int L, int NW, bool F64>
__global__ __launch_bounds__(256) void residue_reconstruct_kernel(int m, int n, int W, ResTab t, ResPlanes pl,
                                                                  const int *__restrict__ E,
                                                                  const int *__restrict__ F, double *__restrict__ C,
                                                                  long long ldc)
{
    residue_reconstruct_at<L, NW, F64>((long long)blockIdx.x * blockDim.x + threadIdx.x, m, n, W, t, pl, E, F, C,
                                       ldc);
}

// --------------------------------------------------------------- moduli --
// Narrow: every m <= 256 coprime to all larger ones taken (256, 255, 253,
// 251, 247, ...; 49 moduli, 341.9 bits).  Wide: primes below 64768 from the
// top (their balanced residues split into two INT8 digits).
static const std::vector<int> &narrow_moduli()
{
    static const std::vector<int> v = [] {
        std::vector<int> out;
        for (int m = 256; m >= 2; --m) {
            bool ok = true;
            for (int x : out) {
                int a = m, b = x;
                while (b) {
                    const int t = a % b;
                    a = b;
                    b = t;
                }
                if (a != 1) {
                    ok = false;
                    break;
                }
            }
            // below 2^4 a narrow modulus adds fewer bits per GEMM than a wide one
            if (ok && m > 16) out.push_back(m);
        }
        return out;
    }();
    return v;
}

static const std::vector<int> &wide_moduli()
{
    static const std::vector<int> v = [] {
        std::vector<int> out;
        for (int p = 64767; p > 257 && out.size() < 192; p -= 2) {
            bool prime = true;
            for (int d = 3; d * d <= p; d += 2)
                if (p % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) out.push_back(p);
        }
        return out;
    }();
    return v;
}

typedef std::vector<uint32_t> Big;

static void big_mul_small(Big &a, uint32_t f)
{
    uint64_t c = 0;
    for (auto &w : a) {
        const uint64_t t = (uint64_t)w * f + c;
        w = (uint32_t)t;
        c = t >> 32;
    }
    if (c) a.push_back((uint32_t)c);
}

static int big_bitlen(const Big &a)
{
    for (int i = (int)a.size() - 1; i >= 0; --i)
        if (a[i]) return 32 * i + 32 - __builtin_clz(a[i]);
    return 0;
}

static long long inv_mod(long long a, long long m)
{
    long long g = m, x = 0, x1 = 1, aa = a % m;
    while (aa) {
        const long long qq = g / aa;
        long long t = g - qq * aa;
        g = aa;
        aa = t;
        t = x - qq * x1;
        x = x1;
        x1 = t;
    }
    return ((x % m) + m) % m;
}

// FP64-MFMA moduli: odd primes below 2^21 from the top (balanced residues
// below 2^20 in magnitude, products below 2^40)
static const std::vector<int> &f64_moduli()
{
    static const std::vector<int> v = [] {
        std::vector<int> out;
        for (int p = (1 << 21) - 1; p > 3 && out.size() < 160; p -= 2) {
            bool prime = true;
            for (int d = 3; d * d <= p; d += 2)
                if (p % d == 0) {
                    prime = false;
                    break;
                }
            if (prime) out.push_back(p);
        }
        return out;
    }();
    return v;
}

static std::vector<int> moduli_of(int n8, int n16, int nf = 0)
{
    if (nf > 0) return std::vector<int>(f64_moduli().begin(), f64_moduli().begin() + nf);
    std::vector<int> v(narrow_moduli().begin(), narrow_moduli().begin() + n8);
    v.insert(v.end(), wide_moduli().begin(), wide_moduli().begin() + n16);
    return v;
}

// Fewest FP64 moduli with M >= 2^need; *bits = bit length of M.
static bool choose_f64_moduli(int need, int &nf, int *bits)
{
    Big M{1u};
    for (nf = 0; nf < (int)f64_moduli().size();) {
        big_mul_small(M, (uint32_t)f64_moduli()[nf++]);
        *bits = big_bitlen(M);
        if (*bits >= need + 1) return true;
    }
    return false;
}

// Cheapest set (n8 + 4 n16 GEMM units) with M >= 2^need; *bits = bit length of M.
static bool choose_moduli(int need, int &n8, int &n16, int *bits)
{
    const auto &a = narrow_moduli();
    const auto &b = wide_moduli();
    std::vector<double> c8(a.size() + 1, 0.0);
    for (size_t i = 0; i < a.size(); ++i) c8[i + 1] = c8[i] + std::log2((double)a[i]);
    int best = INT_MAX;
    double bits16 = 0.0;
    for (size_t j = 0; j <= b.size(); ++j) {
        const double rem = need + 1e-6 - bits16;
        int i8 = -1;
        for (size_t i = 0; i <= a.size(); ++i)
            if (c8[i] >= rem) {
                i8 = (int)i;
                break;
            }
        if (i8 >= 0 && i8 + 4 * (int)j < best) {
            best = i8 + 4 * (int)j;
            n8 = i8;
            n16 = (int)j;
        }
        if (j < b.size()) bits16 += std::log2((double)b[j]);
    }
    if (best == INT_MAX) return false;
    // confirm with exact arithmetic (the log sums carry rounding)
    for (;;) {
        Big M{1u};
        for (int m : moduli_of(n8, n16)) big_mul_small(M, (uint32_t)m);
        *bits = big_bitlen(M);
        if (*bits >= need + 1) return true;
        if (n8 < (int)a.size())
            ++n8;
        else if (n16 < (int)b.size())
            ++n16;
        else
            return false;
    }
}

struct ResTabHost {
    ResTab t{};
    void *dev = nullptr;
};

static std::mutex g_res_mu;
static std::map<std::tuple<int, int, int, int, int>, ResTabHost *> g_res;

static double balanced(long long p, long long m) { return (double)(2 * p >= m ? p - m : p); }

// value of a big integer as 24-bit words (false if it needs more than nw)
static bool big_words24(const Big &b, int nw, double *out)
{
    const int bits = big_bitlen(b);
    if ((bits + 23) / 24 > nw) return false;
    for (int w = 0; w * 24 < bits; ++w) {
        uint32_t v = 0;
        for (int bit = 0; bit < 24; ++bit) {
            const int pos = 24 * w + bit;
            if ((pos >> 5) < (int)b.size() && ((b[pos >> 5] >> (pos & 31)) & 1u)) v |= 1u << bit;
        }
        out[w] = (double)v;
    }
    return true;
}

// device tables of a modulus set with CRT constants of nw 24-bit words (cached)
static ResTabHost *res_table(int n8, int n16, int nf, int nw)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_res_mu);
    const auto key = std::make_tuple(dev, n8, n16, nf, nw);
    auto it = g_res.find(key);
    if (it != g_res.end()) return it->second;

    const std::vector<int> mods = moduli_of(n8, n16, nf);
    const int n = (int)mods.size();
    const int ng8 = (n8 + 2) / 3, G = ng8 + n16 + nf;
    // the groups: narrow triples (3g, 3g+1, 3g+2), then each wide modulus;
    // or each FP64 modulus
    std::vector<long long> P(G);
    std::vector<std::vector<int>> members(G);
    for (int i = 0; i < n8; ++i) members[i / 3].push_back(i);
    for (int j = 0; j < n16; ++j) members[ng8 + j].push_back(n8 + j);
    for (int i = 0; i < nf; ++i) members[i].push_back(i);
    for (int g = 0; g < G; ++g) {
        P[g] = 1;
        for (int i : members[g]) P[g] *= mods[i];
    }
    const size_t o_inv = n, o_p37 = 2 * (size_t)n, o_p37h = o_p37 + (size_t)n8 * RES_NC,
                 o_p37l = o_p37h + (size_t)n16 * RES_NC, o_gP = o_p37l + (size_t)n16 * RES_NC, o_ginv = o_gP + G,
                 o_gfrac = o_ginv + G, o_gE = o_gfrac + G, o_cw = o_gE + 3 * (size_t)ng8, o_M = o_cw + (size_t)G * nw,
                 o_p24 = o_M + nw, o_c32 = o_p24 + (size_t)nf * RES_NC24, total = o_c32 + nf;
    std::vector<double> h(total, 0.0);
    for (int i = 0; i < n; ++i) {
        h[i] = (double)mods[i];
        h[o_inv + i] = 1.0 / (double)mods[i];
    }
    for (int i = 0; i < n8; ++i) {
        long long p = 1 % mods[i];
        for (int c = 0; c < RES_NC; ++c) {
            h[o_p37 + (size_t)i * RES_NC + c] = balanced(p, mods[i]);
            p = (p << 37) % mods[i];   // p < 2^8: no overflow
        }
    }
    for (int j = 0; j < n16; ++j) {
        const long long mj = mods[n8 + j];
        long long p = 1;
        for (int c = 0; c < RES_NC; ++c) {
            const long long b = (long long)balanced(p, mj);   // |b| < 2^15
            const long long lo = ((b + 128) & 255) - 128;
            h[o_p37h + (size_t)j * RES_NC + c] = (double)((b - lo) / 256);
            h[o_p37l + (size_t)j * RES_NC + c] = (double)lo;
            p = (long long)(((unsigned __int128)p << 37) % (uint64_t)mj);
        }
    }
    for (int i = 0; i < nf; ++i) {
        const long long mi = mods[i];
        long long p = 1;
        for (int c = 0; c < RES_NC24; ++c) {
            h[o_p24 + (size_t)i * RES_NC24 + c] = balanced(p, mi);
            p = (p << 24) % mi;   // p < 2^21: no overflow
        }
        h[o_c32 + i] = balanced((1ll << 32) % mi, mi);
    }
    Big M{1u};
    for (int m : mods) big_mul_small(M, (uint32_t)m);
    if (!big_words24(M, nw, &h[o_M])) return nullptr;
    for (int g = 0; g < G; ++g) {
        const long long Pg = P[g];
        Big Mg{1u};
        long long rem = 1 % Pg;
        for (int h2 = 0; h2 < G; ++h2)
            if (h2 != g) {
                big_mul_small(Mg, (uint32_t)P[h2]);
                rem = (long long)((unsigned __int128)rem * (uint64_t)(P[h2] % Pg) % (uint64_t)Pg);
            }
        const long long yg = inv_mod(rem, Pg);
        big_mul_small(Mg, (uint32_t)yg);
        if (!big_words24(Mg, nw, &h[o_cw + (size_t)g * nw])) return nullptr;
        h[o_gP + g] = (double)Pg;
        h[o_ginv + g] = 1.0 / (double)Pg;
        h[o_gfrac + g] = (double)yg / (double)Pg;
        if (g < ng8)
            for (size_t x = 0; x < members[g].size(); ++x) {
                const long long mx = mods[members[g][x]], Px = Pg / mx;
                h[o_gE + 3 * (size_t)g + x] = (double)(Px * inv_mod(Px % mx, mx));
            }
    }
    double *d = nullptr;
    if (hipMalloc(&d, total * sizeof(double)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), total * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    ResTabHost *r = new ResTabHost();
    r->dev = d;
    r->t.n8 = n8;
    r->t.n16 = n16;
    r->t.ng8 = ng8;
    r->t.nw = nw;
    r->t.md = d;
    r->t.inv = d + o_inv;
    r->t.pw37 = d + o_p37;
    r->t.pw37h = d + o_p37h;
    r->t.pw37l = d + o_p37l;
    r->t.gP = d + o_gP;
    r->t.ginv = d + o_ginv;
    r->t.gfrac = d + o_gfrac;
    r->t.gE = d + o_gE;
    r->t.cw = d + o_cw;
    r->t.Mw = d + o_M;
    r->t.nf = nf;
    r->t.pw24 = d + o_p24;
    r->t.c32 = d + o_c32;
    g_res[key] = r;
    return r;
}

// ------------------------------------------------------------ dispatch --
// fixed-point words the residue kernels build (|N| < 2^(W-2) plus a sign bit)
template <int L>
static int nwu_bucket(int W)
{
    constexpr int NW = FixedWords<L>::NW;
    const int need = (W + 31) / 32;
    const int lo = (L == 2) ? 4 : (L == 4) ? 8 : (L == 8) ? 12 : 24;
    return std::min(NW, std::max(lo, (int)round_up(need, 4)));
}

// 24-bit CRT words per limb count: M < 2^mbits and |Z| <= M/4, so the
// words of M hold Z in two's complement with the carry out as its sign
static int nw_bucket(int L, int mbits)
{
    const int words = (mbits + 23) / 24;
    const int lo = (L == 2) ? 12 : (L == 4) ? 16 : (L == 8) ? 32 : 56;
    const int step = (L <= 4) ? 4 : 8;
    return (int)round_up(std::max(words, lo), step);
}

template <int L, int NWU>
static void launch_residues(int m, int n, int k, int mp, int np, int kp, const double *A, long long lda,
                            const double *B, long long ldb, const int *E, const int *F, int W, const ResTab &t,
                            int8_t *A8, int8_t *A16, int8_t *B8, int8_t *B16, hipStream_t s)
{
    constexpr int EPT = ResEPT<L>::value;
    constexpr int TC = 256 / (RES_TK / EPT);
    hipLaunchKernelGGL((residue_a_kernel<L, NWU>), dim3(grid_for((long long)mp * (kp / EPT), 256)), dim3(256), 0, s,
                       m, k, mp, kp, A, lda, E, W, t, A8, A16);
    hipLaunchKernelGGL((residue_b_kernel<L, NWU>), dim3((unsigned)((np / TC) * (kp / RES_TK))), dim3(256), 0, s, k,
                       n, kp, np, B, ldb, F, W, t, B8, B16);
}

template <int L>
static int residues_dispatch(int nwu, int m, int n, int k, int mp, int np, int kp, const double *A, long long lda,
                             const double *B, long long ldb, const int *E, const int *F, int W, const ResTab &t,
                             int8_t *A8, int8_t *A16, int8_t *B8, int8_t *B16, hipStream_t s)
{
#define MDG_RES_CASE(NWU)                                                                                     \
    case NWU:                                                                                                 \
        if constexpr (NWU <= FixedWords<L>::NW) {                                                             \
            launch_residues<L, NWU>(m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, A8, A16, B8, B16, s); \
            return MDG_SUCCESS;                                                                               \
        }                                                                                                     \
        break;
    switch (nwu) {
        MDG_RES_CASE(4)
        MDG_RES_CASE(8)
        MDG_RES_CASE(12)
        MDG_RES_CASE(16)
        MDG_RES_CASE(20)
        MDG_RES_CASE(24)
        MDG_RES_CASE(28)
        MDG_RES_CASE(32)
        MDG_RES_CASE(36)
    }
#undef MDG_RES_CASE
    set_error("no residue kernel for %d fixed-point words", nwu);
    return MDG_ERROR_RANGE;
}

template <int L, int NW>
static void launch_reconstruct(bool f64, int m, int n, int W, const ResTab &t, const ResPlanes &pl, const int *E,
                               const int *F, double *C, long long ldc, hipStream_t s)
{
    if (f64)
        hipLaunchKernelGGL((residue_reconstruct_kernel<L, NW, true>), dim3(grid_for((long long)m * n, 256)),
                           dim3(256), 0, s, m, n, W, t, pl, E, F, C, ldc);
    else
        hipLaunchKernelGGL((residue_reconstruct_kernel<L, NW, false>), dim3(grid_for((long long)m * n, 256)),
                           dim3(256), 0, s, m, n, W, t, pl, E, F, C, ldc);
}

template <int L>
static int reconstruct_dispatch(bool f64, int nw, int m, int n, int W, const ResTab &t, const ResPlanes &pl,
                                const int *E, const int *F, double *C, long long ldc, hipStream_t s)
{
#define MDG_REC_CASE(NW)                                                 \
    case NW:                                                             \
        launch_reconstruct<L, NW>(f64, m, n, W, t, pl, E, F, C, ldc, s); \
        return MDG_SUCCESS;
    if constexpr (L == 2) {
        switch (nw) { MDG_REC_CASE(12) MDG_REC_CASE(16) MDG_REC_CASE(20) MDG_REC_CASE(24) }
    } else if constexpr (L == 4) {
        switch (nw) {
            MDG_REC_CASE(16) MDG_REC_CASE(20) MDG_REC_CASE(24) MDG_REC_CASE(28) MDG_REC_CASE(32) MDG_REC_CASE(36)
        }
    } else if constexpr (L == 8) {
        switch (nw) { MDG_REC_CASE(32) MDG_REC_CASE(40) MDG_REC_CASE(48) MDG_REC_CASE(56) }
    } else {
        switch (nw) {
            MDG_REC_CASE(56) MDG_REC_CASE(64) MDG_REC_CASE(72) MDG_REC_CASE(80) MDG_REC_CASE(88) MDG_REC_CASE(96)
            MDG_REC_CASE(104)
        }
    }
#undef MDG_REC_CASE
    set_error("no reconstruction kernel for %d CRT words at L=%d", nw, L);
    return MDG_ERROR_RANGE;
}

// one strided-batched INT8 GEMM per K-chunk of the depth; chunk q of batch
// entry b lands in plane b * nck + q (plane stride ps)
static int res_batched(int batch, int m, int n, int depth, const int8_t *Bp, long long ldb, long long sb,
                       const int8_t *Ap, long long lda, long long sa, int *D, long long ps, int nck, hipStream_t s,
                       int *engine)
{
    for (int q = 0; q * (long long)RES_KC < depth; ++q) {
        const int k0 = q * RES_KC;
        const int kc = std::min(RES_KC, depth - k0);
        const int rc = i8_gemm_tn(batch, m, n, kc, Bp + k0, ldb, sb, Ap + k0, lda, sa, D + (long long)q * ps, n,
                                  ps * nck, s, engine);
        if (rc) return rc;
    }
    return MDG_SUCCESS;
}

template <int L>
static int run_residue_gemm(int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                            double *C, long long ldc, int target_bits, int fixed_spread, mdg_gemm_info *info,
                            hipStream_t s)
{
    StageTimer tm(s, profiling_on());
    tm.mark(0);
    MDG_REQUIRE(k <= (1 << 22), MDG_ERROR_RANGE, "the residue GEMM supports k <= 2^22 (k=%d)", k);
    int *E = nullptr, *F = nullptr;
    int spread_a = 0, spread_b = 0;
    int rc = scan_scales(L, m, n, k, A, lda, B, ldb, fixed_spread, s, &E, &F, &spread_a, &spread_b);
    if (rc) return rc;
    tm.mark(1);

    int lgL = 0;
    while ((1 << lgL) < L) ++lgL;
    int lgk = 0;
    while ((1ll << lgk) < k) ++lgk;
    const int W = target_bits + spread_a + spread_b + 3 + lgL;
    MDG_REQUIRE(W <= 32 * FixedWords<L>::NW - 1, MDG_ERROR_RANGE,
                "needs %d fixed-point bits, more than the %d this limb count holds", W, 32 * FixedWords<L>::NW - 1);
    const int need = 2 * W - 1 + lgk;   // M >= 2^need > 4 k 2^(2W-4) >= 4 max|Z|
    int n8 = 0, n16 = 0, mbits = 0;
    MDG_REQUIRE(choose_moduli(need, n8, n16, &mbits), MDG_ERROR_RANGE, "no modulus set reaches %d bits", need);
    const int nwu = nwu_bucket<L>(W);
    const int nw = nw_bucket(L, mbits);
    ResTabHost *th = res_table(n8, n16, 0, nw);
    MDG_REQUIRE(th, MDG_ERROR_HIP, "could not build the modulus tables (%d + %d moduli)", n8, n16);
    const ResTab &t = th->t;

    const int kp = (int)round_up(k, 128), mp = (int)round_up(m, 128), np = (int)round_up(n, 128);
    const int nck1 = (kp + RES_KC - 1) / RES_KC, nck2 = (2 * kp + RES_KC - 1) / RES_KC;
    const long long ps = plane_stride((long long)m * n, 4);
    const size_t a8 = (size_t)n8 * mp * kp, a16 = (size_t)n16 * mp * 2 * kp;
    const size_t b8 = (size_t)n8 * np * kp, b16 = (size_t)n16 * np * 3 * kp;
    const size_t nplanes = (size_t)n8 * nck1 + (size_t)n16 * (2 * nck1 + nck2);
    int8_t *Ares = (int8_t *)scratch(s, a8 + a16 + 256, 2);
    int8_t *Bres = (int8_t *)scratch(s, b8 + b16 + 256, 3);
    int *D = (int *)scratch(s, nplanes * ps * 4, 4);
    MDG_REQUIRE(Ares && Bres && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    int8_t *A8 = Ares, *A16 = Ares + a8, *B8 = Bres, *B16 = Bres + b8;
    int *D8 = D, *Dhh = D + (size_t)n8 * nck1 * ps, *Dll = Dhh + (size_t)n16 * nck1 * ps,
        *Dx = Dll + (size_t)n16 * nck1 * ps;

    rc = residues_dispatch<L>(nwu, m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, A8, A16, B8, B16, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("residue kernels");
    tm.mark(2);
    int engine = 1, wide_engine = 1;
    {
        const long long sa8 = (long long)mp * kp, sb8 = (long long)np * kp;
        const long long sa16 = (long long)mp * 2 * kp, sb16 = (long long)np * 3 * kp;
        if (n8 && (rc = res_batched(n8, m, n, kp, B8, kp, sb8, A8, kp, sa8, D8, ps, nck1, s, &engine))) return rc;
        if (n16) {
            // hi*hi, lo*lo, and (hi, lo) . (lo, hi) of depth 2kp
            if ((rc = res_batched(n16, m, n, kp, B16, 3ll * kp, sb16, A16, 2ll * kp, sa16, Dhh, ps, nck1, s,
                                  &wide_engine)))
                return rc;
            if ((rc = res_batched(n16, m, n, kp, B16 + kp, 3ll * kp, sb16, A16 + kp, 2ll * kp, sa16, Dll, ps, nck1,
                                  s, &wide_engine)))
                return rc;
            if ((rc = res_batched(n16, m, n, 2 * kp, B16 + kp, 3ll * kp, sb16, A16, 2ll * kp, sa16, Dx, ps, nck2, s,
                                  &wide_engine)))
                return rc;
        }
    }
    tm.mark(3);
    const ResPlanes pl{D8, Dhh, Dll, Dx, nullptr, ps, nck1, nck2, (nck1 > 1 || kp > 8192) ? 1 : 0};
    rc = reconstruct_dispatch<L>(false, nw, m, n, W, t, pl, E, F, C, ldc, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("residue_reconstruct_kernel");
    tm.mark(4);
    float stage[4] = {0.f, 0.f, 0.f, 0.f};
    if (tm.on) MDG_HIP_CHECK(tm.elapsed(stage));
    if (info) {
        info->method = MDG_METHOD_INT8_RESIDUE;
        info->slices = n8 + n16;
        info->slice_bits = 8;
        info->diagonals = n16;
        info->pair_gemms = n8 + 4ll * n16;
        info->total_bits = W;
        info->spread_a = spread_a;
        info->spread_b = spread_b;
        info->flushed = (nck1 > 1 || nck2 > 1) ? 1 : 0;
        info->backend = engine;   // engine of the narrow-modulus batch, the largest
        for (int i = 0; i < 4; ++i) info->stage_ms[i] = stage[i];
    }
    return MDG_SUCCESS;
}

// ------------------------------------------------ FP64-MFMA residue GEMM --
template <int L, int NWU>
static void launch_f64_residues(int m, int n, int k, int mp, int np, int kp, const double *A, long long lda,
                                const double *B, long long ldb, const int *E, const int *F, int W, const ResTab &t,
                                int32_t *RA, int32_t *RB, hipStream_t s)
{
    hipLaunchKernelGGL((f64_residue_a_kernel<L, NWU>), dim3(grid_for((long long)mp * kp, 256)), dim3(256), 0, s, m,
                       k, mp, kp, A, lda, E, W, t, RA);
    hipLaunchKernelGGL((f64_residue_b_kernel<L, NWU>), dim3((unsigned)((np / 4) * (kp / RES_TK))), dim3(256), 0, s,
                       k, n, kp, np, B, ldb, F, W, t, RB);
}

template <int L>
static int f64_residues_dispatch(int nwu, int m, int n, int k, int mp, int np, int kp, const double *A,
                                 long long lda, const double *B, long long ldb, const int *E, const int *F, int W,
                                 const ResTab &t, int32_t *RA, int32_t *RB, hipStream_t s)
{
#define MDG_RES_CASE(NWU)                                                                               \
    case NWU:                                                                                           \
        if constexpr (NWU <= FixedWords<L>::NW) {                                                       \
            launch_f64_residues<L, NWU>(m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, RA, RB, s); \
            return MDG_SUCCESS;                                                                         \
        }                                                                                               \
        break;
    switch (nwu) {
        MDG_RES_CASE(4)
        MDG_RES_CASE(8)
        MDG_RES_CASE(12)
        MDG_RES_CASE(16)
        MDG_RES_CASE(20)
        MDG_RES_CASE(24)
        MDG_RES_CASE(28)
        MDG_RES_CASE(32)
        MDG_RES_CASE(36)
    }
#undef MDG_RES_CASE
    set_error("no residue kernel for %d fixed-point words", nwu);
    return MDG_ERROR_RANGE;
}

// Residues modulo odd primes below 2^21, one INT32 plane per modulus and
// operand, products on the FP64 MFMA kernel (exact FP64 partial sums moved
// into INT64 every 8192 k: 8192 products of magnitude below 2^40 stay below
// 2^53), then the same CRT reconstruction with one group per modulus.
template <int L>
static int run_residue_f64(int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                           double *C, long long ldc, int target_bits, int fixed_spread, mdg_gemm_info *info,
                           hipStream_t s)
{
    StageTimer tm(s, profiling_on());
    tm.mark(0);
    MDG_REQUIRE(k <= (1 << 22), MDG_ERROR_RANGE, "the residue GEMM supports k <= 2^22 (k=%d)", k);
    int *E = nullptr, *F = nullptr;
    int spread_a = 0, spread_b = 0;
    int rc = scan_scales(L, m, n, k, A, lda, B, ldb, fixed_spread, s, &E, &F, &spread_a, &spread_b);
    if (rc) return rc;
    tm.mark(1);

    int lgL = 0;
    while ((1 << lgL) < L) ++lgL;
    int lgk = 0;
    while ((1ll << lgk) < k) ++lgk;
    const int W = target_bits + spread_a + spread_b + 3 + lgL;
    MDG_REQUIRE(W <= 32 * FixedWords<L>::NW - 1, MDG_ERROR_RANGE,
                "needs %d fixed-point bits, more than the %d this limb count holds", W, 32 * FixedWords<L>::NW - 1);
    const int need = 2 * W - 1 + lgk;
    int nf = 0, mbits = 0;
    MDG_REQUIRE(choose_f64_moduli(need, nf, &mbits), MDG_ERROR_RANGE, "no FP64 modulus set reaches %d bits", need);
    const int nwu = nwu_bucket<L>(W);
    const int nw = nw_bucket(L, mbits);
    ResTabHost *th = res_table(0, 0, nf, nw);
    MDG_REQUIRE(th, MDG_ERROR_HIP, "could not build the modulus tables (%d FP64 moduli)", nf);
    const ResTab &t = th->t;

    const int kp = (int)round_up(k, 128), mp = (int)round_up(m, 128), np = (int)round_up(n, 128);
    const long long ps = plane_stride((long long)m * n, 8);
    int32_t *RA = (int32_t *)scratch(s, (size_t)nf * mp * kp * 4 + 256, 2);
    int32_t *RB = (int32_t *)scratch(s, (size_t)nf * np * kp * 4 + 256, 3);
    long long *D = (long long *)scratch(s, (size_t)nf * ps * 8, 4);
    MDG_REQUIRE(RA && RB && D, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    rc = f64_residues_dispatch<L>(nwu, m, n, k, mp, np, kp, A, lda, B, ldb, E, F, W, t, RA, RB, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("FP64 residue kernels");
    tm.mark(2);
    for (int i = 0; i < nf; ++i) {
        rc = f64_gemm_i32(RA + (size_t)i * mp * kp, kp, RB + (size_t)i * np * kp, kp, m, n, kp, 8192, D + i * ps, s);
        if (rc) return rc;
    }
    tm.mark(3);
    const ResPlanes pl{nullptr, nullptr, nullptr, nullptr, D, ps, 1, 1, 0};
    rc = reconstruct_dispatch<L>(true, nw, m, n, W, t, pl, E, F, C, ldc, s);
    if (rc) return rc;
    MDG_LAUNCH_CHECK("residue_reconstruct_kernel");
    tm.mark(4);
    float stage[4] = {0.f, 0.f, 0.f, 0.f};
    if (tm.on) MDG_HIP_CHECK(tm.elapsed(stage));
    if (info) {
        info->method = MDG_METHOD_F64_RESIDUE;
        info->slices = nf;
        info->slice_bits = 21;
        info->diagonals = 0;
        info->pair_gemms = nf;
        info->total_bits = W;
        info->spread_a = spread_a;
        info->spread_b = spread_b;
        info->flushed = 1;
        info->backend = 2;
        for (int i = 0; i < 4; ++i) info->stage_ms[i] = stage[i];
    }
    return MDG_SUCCESS;
}

// ------------------------------------------------ fused batch pipeline --
// Many INT8 residue GEMMs of uneven shape as three launches per limb count
// (residues of A, of B, reconstruction) plus one launch of the INT8 MFMA
// tiles of every residue plane of every problem.  Each problem gets the
// fixed-point width the single-problem path picks for it with the batch's
// fixed spread; the results are bit-identical to a loop of mdg_gemm calls.
// The exponent scans (E, F per problem) come from the caller's pipeline.

// one problem, as the kernels see it; residue planes as in run_residue_gemm.
// Every problem of one limb count uses that limb count's modulus set (the
// one its deepest problem needs: Z is the same exact integer whatever valid
// set computes it, so the results stay bit-identical to single calls).
// Result planes, reduced to [0, m) by the GEMM epilogue: narrow i at
// P8 + i ps (bytes); wide j: hh at P16 + j ps, ll at (n16 + j) ps, cross at
// (2 n16 + j) ps (16-bit words).
struct RProb {
    int L, m, n, k, mp, np, kp, W, n8, n16;
    long long lda, ldb, ldc, ps;
    const double *A, *B;
    double *C;
    const int *E, *F;
    int8_t *A8, *A16, *B8, *B16;
    uint8_t *P8;
    uint16_t *P16;
    const double *md, *inv;   // the moduli of the problem's set and their reciprocals
};

struct REnt {
    int prob, x, y, z;
};

// the limb count's modulus tables come as a kernel argument (read with
// scalar loads, as in the single-problem kernels)
template <int L, int NWU>
__global__ __launch_bounds__(256) void b_residue_a_kernel(const RProb *__restrict__ P, const REnt *__restrict__ T,
                                                          ResTab t)
{
    const REnt e = T[blockIdx.x];
    const RProb p = P[e.prob];
    residue_a_at<L, NWU>((long long)e.x * 256 + threadIdx.x, p.m, p.k, p.mp, p.kp, p.A, p.lda, p.E, p.W, t, p.A8,
                         p.A16);
}

template <int L, int NWU>
__global__ __launch_bounds__(256) void b_residue_b_kernel(const RProb *__restrict__ P, const REnt *__restrict__ T,
                                                          ResTab t)
{
    const REnt e = T[blockIdx.x];
    const RProb p = P[e.prob];
    residue_b_at<L, NWU>(e.x, p.k, p.n, p.kp, p.np, p.B, p.ldb, p.F, p.W, t, p.B8, p.B16);
}

// entry (prob, tile row, tile column, plane): a TM x TM tile (i8tile.hpp);
// plane z < n8 is narrow modulus z; then wide modulus j = (z - n8) / 3 with
// v = (z - n8) % 3: 0 hi*hi, 1 lo*lo, 2 (hi, lo) . (lo, hi) of depth 2kp.
// The epilogue stores each exact INT32 sum reduced modulo its modulus.
template <int TM>
__global__ __launch_bounds__(256) void b_residue_gemm_kernel(const RProb *__restrict__ P, const REnt *__restrict__ T)
{
    __shared__ __attribute__((aligned(16))) int8_t As[TM * I8T_LDR];
    __shared__ __attribute__((aligned(16))) int8_t Bs[TM * I8T_LDR];
    const REnt e = T[blockIdx.x];
    const RProb p = P[e.prob];
    const long long kp = p.kp;
    const int8_t *Ab, *Bb;
    long long lda, ldb;
    int Kd, mi;
    if (e.z < p.n8) {
        Ab = p.A8 + (long long)e.z * p.mp * kp;
        Bb = p.B8 + (long long)e.z * p.np * kp;
        lda = ldb = kp;
        Kd = (int)kp;
        mi = e.z;
    } else {
        const int j = (e.z - p.n8) / 3, v = (e.z - p.n8) % 3;
        lda = 2 * kp;
        ldb = 3 * kp;
        Ab = p.A16 + (long long)j * p.mp * lda + (v == 1 ? kp : 0);
        Bb = p.B16 + (long long)j * p.np * ldb + (v == 0 ? 0 : kp);
        Kd = (int)(v == 2 ? 2 * kp : kp);
        mi = p.n8 + j;
    }
    i4t acc[I8Tile<TM>::IM][I8Tile<TM>::IM];
    i8_tile<TM, I8Tile<TM>::RING>(Ab + (long long)e.x * TM * lda, lda, Bb + (long long)e.y * TM * ldb, ldb, Kd, As, Bs, acc);
    const double md = p.md[mi], inv = p.inv[mi];
    const long long row0 = (long long)e.x * TM, col0 = (long long)e.y * TM;
    const int n = p.n;
    if (e.z < p.n8) {
        uint8_t *plane = p.P8 + (long long)e.z * p.ps;
        i8_tile_each<TM>(acc, row0, col0, p.m, n, [&](long long r, long long c, int v) {
            plane[r * n + c] = (uint8_t)(int)mod_pos((double)v, md, inv);
        });
    } else {
        const int j = (e.z - p.n8) / 3, v3 = (e.z - p.n8) % 3;
        uint16_t *plane = p.P16 + (long long)(v3 * p.n16 + j) * p.ps;
        i8_tile_each<TM>(acc, row0, col0, p.m, n, [&](long long r, long long c, int v) {
            plane[r * n + c] = (uint16_t)(int)mod_pos((double)v, md, inv);
        });
    }
}

template <int L, int NW>
__global__ __launch_bounds__(256) void b_residue_reconstruct_kernel(const RProb *__restrict__ P,
                                                                    const REnt *__restrict__ T, ResTab t)
{
    const REnt e = T[blockIdx.x];
    const RProb p = P[e.prob];
    ResPlanes pl{nullptr, nullptr, nullptr, nullptr, nullptr, p.ps, 1, 1, 0};
    pl.P8 = p.P8;
    pl.P16 = p.P16;
    residue_reconstruct_at<L, NW, false, true>((long long)e.x * 256 + threadIdx.x, p.m, p.n, p.W, t, pl, p.E, p.F,
                                               p.C, p.ldc);
}

struct ResBatch {
    struct St {
        int kind;   // 0 residues of A, 1 of B, 2 GEMM tiles, 3 reconstruction
        int L, nwu, nw;
        long long first, count;
        int tm = 128;   // GEMM tiles: rows and columns
    };
    std::vector<RProb> probs;
    std::vector<REnt> ents;
    std::vector<St> stages;
    std::vector<St> gemm_cands;   // the GEMM stage with 128- and 64-row tiles
    int gemm_pick = 0;
    ResTab tab[17] = {};          // per limb count
    RProb *dprobs = nullptr;
    REnt *dents = nullptr;
    void *arena = nullptr;
    ~ResBatch()
    {
        (void)hipFree(dprobs);
        (void)hipFree(dents);
        (void)hipFree(arena);
    }
};

static int fixed_words(int L)
{
    return L == 2 ? FixedWords<2>::NW : L == 4 ? FixedWords<4>::NW : L == 8 ? FixedWords<8>::NW : FixedWords<16>::NW;
}
static int res_ept(int L) { return L <= 4 ? 4 : (L == 8 ? 2 : 1); }
static int nwu_for(int L, int W)
{
    switch (L) {
    case 2: return nwu_bucket<2>(W);
    case 4: return nwu_bucket<4>(W);
    case 8: return nwu_bucket<8>(W);
    default: return nwu_bucket<16>(W);
    }
}

// E[i], F[i]: device scale exponents of problem i (filled by the caller's
// scans before residue_batch_run).  MDG_ERROR_NOT_SUPPORTED when a problem
// is too deep for one INT32 plane per GEMM (2 kp > RES_KC).
int residue_batch_build(int count, const mdg_problem *pr, int *const *E, int *const *F, int target_bits, int spread,
                        ResBatch **out)
{
    *out = nullptr;
    ResBatch *rb = new ResBatch();
    auto fail = [&](int rc) {
        delete rb;
        return rc;
    };
    // per limb count: fixed-point width and the deepest problem's need
    int Wl[17] = {0}, needl[17] = {0}, n8l[17] = {0}, n16l[17] = {0}, nwl[17] = {0};
    for (int i = 0; i < count; ++i) {
        const mdg_problem &q = pr[i];
        if (q.m == 0 || q.n == 0 || q.k == 0) continue;
        const int L = q.L;
        int lgL = 0, lgk = 0;
        while ((1 << lgL) < L) ++lgL;
        while ((1ll << lgk) < q.k) ++lgk;
        const int tb = target_bits > 0 ? target_bits : 53 * L;
        const int W = tb + 2 * spread + 3 + lgL;
        if (W > 32 * fixed_words(L) - 1) {
            set_error("problem %d needs %d fixed-point bits", i, W);
            return fail(MDG_ERROR_RANGE);
        }
        if (2 * round_up(q.k, 128) > RES_KC) {
            set_error("problem %d: k=%d too deep for the fused residue batch", i, q.k);
            return fail(MDG_ERROR_NOT_SUPPORTED);
        }
        Wl[L] = W;
        needl[L] = std::max(needl[L], 2 * W - 1 + lgk);
    }
    for (int L : {2, 4, 8, 16}) {
        if (!needl[L]) continue;
        int mbits = 0;
        if (!choose_moduli(needl[L], n8l[L], n16l[L], &mbits)) {
            set_error("no modulus set reaches %d bits", needl[L]);
            return fail(MDG_ERROR_RANGE);
        }
        nwl[L] = nw_bucket(L, mbits);
        ResTabHost *th = res_table(n8l[L], n16l[L], 0, nwl[L]);
        if (!th) {
            set_error("could not build the modulus tables (%d + %d moduli)", n8l[L], n16l[L]);
            return fail(MDG_ERROR_HIP);
        }
        rb->tab[L] = th->t;
    }
    size_t bytes = 0;
    auto take = [&](size_t b) {
        const size_t off = bytes;
        bytes += (b + 255) / 256 * 256;
        return off;
    };
    struct Offs {
        size_t a8, a16, b8, b16, p8, p16;
    };
    std::vector<Offs> offs(count);
    rb->probs.resize(count);
    for (int i = 0; i < count; ++i) {
        const mdg_problem &q = pr[i];
        RProb &b = rb->probs[i];
        std::memset(&b, 0, sizeof(b));
        if (q.m == 0 || q.n == 0 || q.k == 0) continue;
        const int L = q.L;
        b.L = L;
        b.m = q.m;
        b.n = q.n;
        b.k = q.k;
        b.kp = (int)round_up(q.k, 128);
        b.mp = (int)round_up(q.m, 128);
        b.np = (int)round_up(q.n, 128);
        b.W = Wl[L];
        b.n8 = n8l[L];
        b.n16 = n16l[L];
        b.lda = q.lda;
        b.ldb = q.ldb;
        b.ldc = q.ldc;
        b.ps = (long long)q.m * q.n;
        b.A = q.A;
        b.B = q.B;
        b.C = q.C;
        b.E = E[i];
        b.F = F[i];
        b.md = rb->tab[L].md;
        b.inv = rb->tab[L].inv;
        offs[i].a8 = take((size_t)b.n8 * b.mp * b.kp);
        offs[i].a16 = take((size_t)b.n16 * b.mp * 2 * b.kp);
        offs[i].b8 = take((size_t)b.n8 * b.np * b.kp);
        offs[i].b16 = take((size_t)b.n16 * b.np * 3 * b.kp);
        offs[i].p8 = take((size_t)b.n8 * b.ps);
        offs[i].p16 = take((size_t)3 * b.n16 * b.ps * 2);
    }
    if (bytes && hipMalloc(&rb->arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("could not allocate %zu bytes for the residue batch", bytes);
        return fail(MDG_ERROR_OUT_OF_MEMORY);
    }
    char *base = (char *)rb->arena;
    for (int i = 0; i < count; ++i) {
        RProb &b = rb->probs[i];
        if (!b.m) continue;
        b.A8 = (int8_t *)(base + offs[i].a8);
        b.A16 = (int8_t *)(base + offs[i].a16);
        b.B8 = (int8_t *)(base + offs[i].b8);
        b.B16 = (int8_t *)(base + offs[i].b16);
        b.P8 = (uint8_t *)(base + offs[i].p8);
        b.P16 = (uint16_t *)(base + offs[i].p16);
    }
    // entry tables
    auto &ents = rb->ents;
    auto stage = [&](int kind, int L, auto gen) {
        ResBatch::St st{kind, L, L ? nwu_for(L, Wl[L]) : 0, L ? nwl[L] : 0, (long long)ents.size(), 0};
        for (int i = 0; i < count; ++i) {
            const RProb &b = rb->probs[i];
            if (!b.m || (L && b.L != L)) continue;
            gen(i, b);
        }
        st.count = (long long)ents.size() - st.first;
        if (st.count) rb->stages.push_back(st);
    };
    const int Ls[4] = {2, 4, 8, 16};
    for (int L : Ls)
        stage(0, L, [&](int i, const RProb &b) {
            const long long nb = ((long long)b.mp * (b.kp / res_ept(L)) + 255) / 256;
            for (long long x = 0; x < nb; ++x) ents.push_back(REnt{i, (int)x, 0, 0});
        });
    for (int L : Ls)
        stage(1, L, [&](int i, const RProb &b) {
            const int tc = 256 / (RES_TK / res_ept(L));
            const long long nb = (long long)(b.np / tc) * (b.kp / RES_TK);
            for (long long x = 0; x < nb; ++x) ents.push_back(REnt{i, (int)x, 0, 0});
        });
    // GEMM tiles of every plane of every problem in one launch.  Workgroup
    // b runs on XCD b % 8, and each XCD has its own L2: the tiles of one
    // (problem, plane) go to one XCD, so the tiles that re-read a residue
    // row block or column block find it in that XCD's L2.  Planes are dealt
    // to the XCD with the least work so far (deep cross planes first), and
    // the eight lists are interleaved.  Built for 128- and 64-row tiles;
    // the first call times both.
    auto gemm_stage = [&](int tm) {
        struct Grp {
            int prob, z;
            long long work;
        };
        std::vector<Grp> groups;
        for (int i = 0; i < count; ++i) {
            const RProb &b = rb->probs[i];
            if (!b.m) continue;
            const long long tiles = (long long)(b.mp / tm) * (b.np / tm);
            for (int z = 0; z < b.n8 + 3 * b.n16; ++z) {
                const bool cross = z >= b.n8 && (z - b.n8) % 3 == 2;
                groups.push_back(Grp{i, z, tiles * b.kp * (cross ? 2 : 1)});
            }
        }
        std::stable_sort(groups.begin(), groups.end(), [](const Grp &a, const Grp &b) { return a.work > b.work; });
        constexpr int NX = 8;
        std::vector<REnt> lists[NX];
        long long load[NX] = {0};
        for (const Grp &g : groups) {
            int x = 0;
            for (int c = 1; c < NX; ++c)
                if (load[c] < load[x]) x = c;
            load[x] += g.work;
            const RProb &b = rb->probs[g.prob];
            for (int tx = 0; tx < b.mp / tm; ++tx)
                for (int ty = 0; ty < b.np / tm; ++ty) lists[x].push_back(REnt{g.prob, tx, ty, g.z});
        }
        ResBatch::St st{2, 0, 0, 0, (long long)ents.size(), 0};
        st.tm = tm;
        size_t longest = 0;
        for (const auto &l : lists) longest = std::max(longest, l.size());
        for (size_t i = 0; i < longest; ++i)
            for (int x = 0; x < NX; ++x)
                if (i < lists[x].size()) ents.push_back(lists[x][i]);
        st.count = (long long)ents.size() - st.first;
        return st;
    };
    {
        const ResBatch::St s128 = gemm_stage(128), s64 = gemm_stage(64);
        if (s128.count) {
            rb->gemm_cands = {s128, s64};
            rb->stages.push_back(s128);   // placeholder: run uses gemm_cands[gemm_pick]
        }
    }
    for (int L : Ls)
        stage(3, L, [&](int i, const RProb &b) {
            const long long nb = ((long long)b.m * b.n + 255) / 256;
            for (long long x = 0; x < nb; ++x) ents.push_back(REnt{i, (int)x, 0, 0});
        });
    for (const auto &st : rb->stages)
        if (st.count > 2147483647ll) {
            set_error("residue batch too large for one launch");
            return fail(MDG_ERROR_NOT_SUPPORTED);
        }
    if (hipMalloc(&rb->dprobs, sizeof(RProb) * std::max(1, count)) != hipSuccess ||
        hipMemcpy(rb->dprobs, rb->probs.data(), sizeof(RProb) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&rb->dents, sizeof(REnt) * std::max<size_t>(1, ents.size())) != hipSuccess ||
        hipMemcpy(rb->dents, ents.data(), sizeof(REnt) * ents.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        set_error("could not upload the residue batch plan");
        return fail(MDG_ERROR_HIP);
    }
    *out = rb;
    return MDG_SUCCESS;
}

template <int L>
static int b_res_launch(const ResBatch::St &st, const ResTab &t, const RProb *P, const REnt *T, hipStream_t s)
{
    const dim3 grid((unsigned)st.count), blk(256);
    if (st.kind == 0 || st.kind == 1) {
#define MDG_BRES_CASE(NWU)                                                                  \
    case NWU:                                                                               \
        if constexpr (NWU <= FixedWords<L>::NW) {                                           \
            if (st.kind == 0)                                                               \
                hipLaunchKernelGGL((b_residue_a_kernel<L, NWU>), grid, blk, 0, s, P, T, t); \
            else                                                                            \
                hipLaunchKernelGGL((b_residue_b_kernel<L, NWU>), grid, blk, 0, s, P, T, t); \
            return MDG_SUCCESS;                                                             \
        }                                                                                   \
        break;
        switch (st.nwu) {
            MDG_BRES_CASE(4)
            MDG_BRES_CASE(8)
            MDG_BRES_CASE(12)
            MDG_BRES_CASE(16)
            MDG_BRES_CASE(20)
            MDG_BRES_CASE(24)
            MDG_BRES_CASE(28)
            MDG_BRES_CASE(32)
            MDG_BRES_CASE(36)
        }
#undef MDG_BRES_CASE
        set_error("no residue kernel for %d fixed-point words", st.nwu);
        return MDG_ERROR_RANGE;
    }
#define MDG_BREC_CASE(NW)                                                                  \
    case NW:                                                                               \
        hipLaunchKernelGGL((b_residue_reconstruct_kernel<L, NW>), grid, blk, 0, s, P, T, t); \
        return MDG_SUCCESS;
    if constexpr (L == 2) {
        switch (st.nw) { MDG_BREC_CASE(12) MDG_BREC_CASE(16) MDG_BREC_CASE(20) MDG_BREC_CASE(24) }
    } else if constexpr (L == 4) {
        switch (st.nw) {
            MDG_BREC_CASE(16) MDG_BREC_CASE(20) MDG_BREC_CASE(24) MDG_BREC_CASE(28) MDG_BREC_CASE(32)
            MDG_BREC_CASE(36)
        }
    } else if constexpr (L == 8) {
        switch (st.nw) { MDG_BREC_CASE(32) MDG_BREC_CASE(40) MDG_BREC_CASE(48) MDG_BREC_CASE(56) }
    } else {
        switch (st.nw) {
            MDG_BREC_CASE(56) MDG_BREC_CASE(64) MDG_BREC_CASE(72) MDG_BREC_CASE(80) MDG_BREC_CASE(88)
            MDG_BREC_CASE(96) MDG_BREC_CASE(104)
        }
    }
#undef MDG_BREC_CASE
    set_error("no reconstruction kernel for %d CRT words at L=%d", st.nw, L);
    return MDG_ERROR_RANGE;
}

int residue_batch_run(const ResBatch *rb, hipStream_t s, hipEvent_t g0, hipEvent_t g1)
{
    for (const auto &st : rb->stages) {
        const REnt *T = rb->dents + st.first;
        int rc = MDG_SUCCESS;
        if (st.kind == 2 && g0) MDG_HIP_CHECK(hipEventRecord(g0, s));
        if (st.kind == 2) {
            const ResBatch::St &g = rb->gemm_cands[rb->gemm_pick];
            const REnt *TG = rb->dents + g.first;
            if (g.tm == 64)
                hipLaunchKernelGGL(b_residue_gemm_kernel<64>, dim3((unsigned)g.count), dim3(256), 0, s, rb->dprobs, TG);
            else
                hipLaunchKernelGGL(b_residue_gemm_kernel<128>, dim3((unsigned)g.count), dim3(256), 0, s, rb->dprobs,
                                   TG);
        } else {
            const ResTab &t = rb->tab[st.L];
            switch (st.L) {
            case 2: rc = b_res_launch<2>(st, t, rb->dprobs, T, s); break;
            case 4: rc = b_res_launch<4>(st, t, rb->dprobs, T, s); break;
            case 8: rc = b_res_launch<8>(st, t, rb->dprobs, T, s); break;
            default: rc = b_res_launch<16>(st, t, rb->dprobs, T, s); break;
            }
        }
        if (rc) return rc;
        MDG_LAUNCH_CHECK("residue batch stage");
        if (st.kind == 2 && g1) MDG_HIP_CHECK(hipEventRecord(g1, s));
    }
    return MDG_SUCCESS;
}

int residue_batch_candidates(const ResBatch *rb) { return (int)rb->gemm_cands.size(); }

void residue_batch_select(ResBatch *rb, int c)
{
    if (c >= 0 && c < (int)rb->gemm_cands.size()) rb->gemm_pick = c;
}

void residue_batch_free(ResBatch *rb) { delete rb; }

int residue_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                 double *C, long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info,
                 hipStream_t s)
{
    if (target_bits <= 0) target_bits = 53 * L;
    if (method == MDG_METHOD_F64_RESIDUE) {
        switch (L) {
        case 2: return run_residue_f64<2>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
        case 4: return run_residue_f64<4>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
        case 8: return run_residue_f64<8>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
        case 16: return run_residue_f64<16>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
        }
        set_error("unsupported limb count L=%d", L);
        return MDG_ERROR_INVALID_VALUE;
    }
    switch (L) {
    case 2: return run_residue_gemm<2>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 4: return run_residue_gemm<4>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 8: return run_residue_gemm<8>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    case 16: return run_residue_gemm<16>(m, n, k, A, lda, B, ldb, C, ldc, target_bits, fixed_spread, info, s);
    }
    set_error("unsupported limb count L=%d", L);
    return MDG_ERROR_INVALID_VALUE;
}

}  // namespace mdg
