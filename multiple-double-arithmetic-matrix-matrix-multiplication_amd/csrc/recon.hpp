// recon.hpp -- exact reconstruction of a slice GEMM entry from its diagonal
// sums.
//
// The entry is C = sum_g D_g 2^(sc0 + w (S-1-g)), g = 0 (most significant)
// .. S-1, with exact integer diagonal sums D_g (INT64).  The reference folds
// such parts one at a time with inc_d (C++/TCcuda.cu:876-906); each fold is a
// full renormalisation.  Here the whole sum is first formed exactly as an
// integer in ws-bit word columns (ws = w * gs, gs = max(1, floor(26 / w))
// diagonals per word: three INT8 diagonals per 24-bit word), written
// straight into a register array -- word i holds the group of diagonals
// nt-1-i, so no run-time register indexing and no carries until one final
// pass -- and then folded into L limbs from its leading (2 ws)-bit piece:
// the first L + 1 pieces through fast_renorm, the rest through inc_d.  The
// only rounding is that final fold.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "common.hpp"
#include "md_arith.hpp"

namespace mdg {

// diagonals per word: w = 8 -> 3 (24-bit words), w <= 13 -> 2, else 1; the
// word width ws = w * gs is then >= 14 bits and <= 26
__host__ __device__ inline int diag_group(int w) { return (26 / w) > 3 ? 3 : ((26 / w) > 1 ? 26 / w : 1); }

// words that hold sum_g D_g 2^(w (S-1-g)) for |D_g| < 2^63
inline int diag_words_needed(int S, int w)
{
    const int gs = diag_group(w), ws = w * gs;
    const int nt = (S + gs - 1) / gs;
    return nt + (63 + ws - 1) / ws + 1;
}

constexpr int RECON_BATCH = 16;   // words whose diagonal loads issue together

// out[s] = d[tp - s] for s < K (0 where tp - s < 0), 0 <= tp < NP, tp per
// lane: the reversed array shifted by NP - 1 - tp in log2(NP) stages of
// selects (instead of one select per (s, p) pair)
template <int NP, int K>
__device__ __forceinline__ void top_pieces(const double (&d)[NP], int tp, double (&out)[K])
{
    double r[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) r[i] = d[NP - 1 - i];
    const int sh = NP - 1 - tp;
#pragma unroll
    for (int b = 0; (1 << b) < NP; ++b) {
        const bool on = (sh >> b) & 1;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double nxt = (i + (1 << b) < NP) ? r[i + (1 << b)] : 0.0;
            r[i] = on ? nxt : r[i];
        }
    }
#pragma unroll
    for (int s = 0; s < K; ++s) out[s] = (s < NP) ? r[s] : 0.0;
}

// The fold of |Z| 2^(w pad), given as NW words of ws = w GS bits (z[0]
// least significant), into L limbs, negated when neg: pieces of two words
// (2 ws <= 52 bits, exact) from the leading nonzero one, the first L + 1
// through fast_renorm, the rest through inc_d; sc0 is the scale exponent of
// Z's unit.  Any exact way of forming the words gives the same limbs.
template <int L, int NW, int GS>
__device__ __forceinline__ void fold_words(const long long (&z)[NW], bool neg, int w, int pad, int sc0, double (&y)[L])
{
    const int ws = w * GS;
    // pieces of two words (2 ws <= 52 bits, exact), the leading one first:
    // K pieces carry at least 1 + 2 ws (K - 1) >= 53 L + 20 bits; the first
    // L + 1 go through fast_renorm, the others one at a time through inc_d
    constexpr int NP = (NW + 1) / 2;
    constexpr int WSMIN = GS == 3 ? 24 : (GS == 2 ? 16 : 14);
    constexpr int KMAX = (53 * L + 19) / (2 * WSMIN) + 2;
    const int K = (53 * L + 19) / (2 * ws) + 2;
    double d[NP];
    int tp = -1;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const long long hi = (2 * p + 1 < NW) ? z[2 * p + 1] : 0;
        d[p] = fma((double)hi, ldexp(1.0, ws), (double)z[2 * p]);
        tp = (d[p] != 0.0) ? p : tp;
    }
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = 0.0;
    if (tp < 0) return;
    const int sc = sc0 - w * pad;
    double top[KMAX];
    top_pieces<NP, KMAX>(d, tp, top);
    auto piece = [&](int s) { return ldexp(top[s], 2 * ws * (tp - s) + sc); };
    {
        double f[L + 1];
#pragma unroll
        for (int s = 0; s <= L; ++s) f[s] = (s < K) ? piece(s) : 0.0;
        fast_renorm<L>(f, y);
    }
#pragma unroll
    for (int s = L + 1; s < KMAX; ++s) {
        if (s >= K) break;
        const double v = piece(s);
        if (v != 0.0) md_inc_d<L>(y, v);
    }
    if (neg) {
#pragma unroll
        for (int l = 0; l < L; ++l) y[l] = -y[l];
    }
}

// D(g) returns the diagonal sum g (0 <= g < S; it is called with every index
// of the word layout, clamped into range, so that all loads issue as
// straight-line code) as an INT64, or as an INT32 when every sum fits one
// (an unchunked INT8 plane); GS = diag_group(w); y receives the L limbs
template <int L, int NW, int GS, typename DiagFn>
__device__ __forceinline__ void diag_fold(int S, int w, int sc0, DiagFn D, double (&y)[L])
{
    // INT32 sums: D_g 2^sh with sh <= w (GS-1) stays below 2^47 and a word's
    // GS of them below 2^49, so each sum is added whole to its own word
    // (the carry pass below spreads it); INT64 sums are split at ws bits
    constexpr bool NARROW = std::is_same_v<decltype(D(0)), int>;
    const int ws = w * GS;
    const int nt = (S + GS - 1) / GS, pad = nt * GS - S;
    long long z[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) z[i] = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int t = nt - 1 - i;
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            const int g = t * GS + u;
            // unconditional load (clamped index) and a mask, not a branch:
            // a branch around each load makes every load wait at its join
            const int sh = w * (GS - 1 - u);
            if constexpr (NARROW) {
                const int dg = D(min(max(g, 0), S - 1)) & -(int)(t >= 0 && g < S);
                z[i] += (long long)dg << sh;
            } else {
                const long long on = -(long long)(t >= 0 && g < S);
                const long long dg = D(min(max(g, 0), S - 1)) & on;
                // D_g 2^sh, sh = w (GS-1-u): the bits below ws into word i,
                // the rest (arithmetic shift) into word i+1 -- no overflow
                z[i] += (dg & ((1ll << (ws - sh)) - 1)) << sh;
                if (i + 1 < NW) z[i + 1] += dg >> (ws - sh);
            }
        }
        // the loads of RB words issue together; the barrier keeps the
        // scheduler from hoisting all of them (registers, then occupancy)
        if (i % RECON_BATCH == RECON_BATCH - 1) __builtin_amdgcn_sched_barrier(0);
    }
    // sum z_i 2^(ws i) = Z 2^(w pad), Z = sum_g D_g 2^(w (S-1-g)); carry into
    // ws-bit words, two's complement with the final carry as the sign
    const long long mask = (1ll << ws) - 1;
    long long carry = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const long long v = z[i] + carry;
        z[i] = v & mask;
        carry = v >> ws;
    }
    const bool neg = carry < 0;
    if (neg) {
        long long cy = 1;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const long long v = (mask - z[i]) + cy;
            z[i] = v & mask;
            cy = v >> ws;
        }
    }
    fold_words<L, NW, GS>(z, neg, w, pad, sc0, y);
}

// bits [24 J, 24 J + 24) of the bit string held in words of WB bits, u[0]
// least significant (J, and so every index, a compile-time constant)
template <int NU, int WB, int J>
__device__ __forceinline__ long long bits24_at(const uint32_t (&u)[NU])
{
    constexpr int bit = 24 * J, wi = bit / WB, off = bit % WB;
    const uint32_t lo = (wi < NU) ? u[wi] : 0u;
    const uint32_t hi = (wi + 1 < NU) ? u[wi + 1] : 0u;
    const uint64_t both = (uint64_t)lo | ((uint64_t)hi << WB);
    return (long long)((both >> off) & 0xffffffu);
}

// z[J] = bits24_at<J>(u) for every J, expanded at compile time (a loop
// this long is not always unrolled, and a run-time index would put u in
// scratch memory)
template <int NU, int WB, int... J>
__device__ __forceinline__ void recut24(const uint32_t (&u)[NU], long long (&z)[sizeof...(J)],
                                        std::integer_sequence<int, J...>)
{
    ((z[J] = bits24_at<NU, WB, J>(u)), ...);
}

// Band planes (i8band.hpp): per band q of G diagonals (first diagonal
// g0_q = S - G (nb - q), nb = ceil(S / G)) and K-chunk c, the exact INT64
// P = sum_j D_{g0_q + j} 2^(w (G-1-j)), so that Z = sum_g D_g 2^(w (S-1-g))
// = sum_q (sum_c P_qc) 2^(w G (nb-1-q)): band q lands whole on word nb-1-q of
// a layout of wb = w G-bit words (32 for INT8 bands of 4, 24 for BF16 bands
// of 3).  The planes of band q follow those of bands 0 .. q-1; P(p) returns
// the value of plane p, nch(q) band q's chunk count and nplanes their total
// (CHUNKED false: one plane per band, plane q, P < 2^(31 + w (G-1)) added
// whole; true: each P split at wb bits, the planes walked from the last
// band down with a running base).  The words are carried into two's complement,
// the magnitude taken, and re-cut into the 24-bit words of Z 2^(8 pad),
// pad = 3 ceil(S/3) - S, which fold_words takes for w = 8 -- the same
// pieces, hence the same limbs, as diag_fold over the diagonal planes.
template <int L, int NWB, int G, bool CHUNKED, typename PlaneFn, typename ChunkFn>
__device__ __forceinline__ void band_fold(int S, int sc0, PlaneFn P, ChunkFn nch, int nplanes, double (&y)[L])
{
    constexpr int w = 8, WB = w * G;
    static_assert(WB == 24 || WB == 32, "band words of 24 or 32 bits");
    constexpr int NW24 = (NWB * WB + 16 + 23) / 24;   // 24-bit words of |Z| 2^(8 pad)
    const int nb = (S + G - 1) / G;
    long long z[NWB];
#pragma unroll
    for (int i = 0; i < NWB; ++i) z[i] = 0;
    int base = nplanes;   // first plane of band q + 1 (chunked)
#pragma unroll
    for (int i = 0; i < NWB; ++i) {
        const int q = nb - 1 - i;
        if constexpr (!CHUNKED) {
            // unconditional load (clamped band) and a mask, not a branch
            z[i] += P(max(q, 0)) & -(long long)(q >= 0);
        } else {
            if (q < 0) break;   // uniform: the same S on every lane
            const int nc = nch(q);
            base -= nc;
            for (int c = 0; c < nc; ++c) {
                const long long v = P(base + c);
                z[i] += v & ((1ll << WB) - 1);
                if (i + 1 < NWB) z[i + 1] += v >> WB;
            }
        }
        if (i % RECON_BATCH == RECON_BATCH - 1) __builtin_amdgcn_sched_barrier(0);
    }
    const long long mask = (1ll << WB) - 1;
    long long carry = 0;
#pragma unroll
    for (int i = 0; i < NWB; ++i) {
        const long long v = z[i] + carry;
        z[i] = v & mask;
        carry = v >> WB;
    }
    const bool neg = carry < 0;
    if (neg) {
        long long cy = 1;
#pragma unroll
        for (int i = 0; i < NWB; ++i) {
            const long long v = (mask - z[i]) + cy;
            z[i] = v & mask;
            cy = v >> WB;
        }
    }
    // |Z| 2^(8 pad) in 24-bit words: shift the WB-bit words left by sh bits
    // (sh < WB), then cut the bit string at multiples of 24
    const int pad = 3 * ((S + 2) / 3) - S, sh = 8 * pad;
    uint32_t u[NWB + 1];
#pragma unroll
    for (int i = 0; i <= NWB; ++i) {
        const uint32_t cur = (i < NWB) ? (uint32_t)z[i] : 0u, prev = (i > 0) ? (uint32_t)z[i - 1] : 0u;
        if constexpr (WB == 32)
            u[i] = sh ? ((cur << sh) | (prev >> (32 - sh))) : cur;
        else
            u[i] = ((cur << sh) | (prev >> (24 - sh))) & 0xffffffu;
    }
    long long z24[NW24];
    recut24<NWB + 1, WB>(u, z24, std::make_integer_sequence<int, NW24>{});
    fold_words<L, NW24, 3>(z24, neg, w, pad, sc0, y);
}

// band-fold word count NWB (WB = 8 G bits) for S slices: the bands, then
// room for the top band's 64-bit sums and the sign; fn is called with
// std::integral_constant<int, NWB> of the smallest fitting instantiation
template <int L, int G, typename Fn>
inline int with_band_words(int S, Fn fn)
{
    const int need = (S + G - 1) / G + (64 + 8 * G - 1) / (8 * G) + 1;
    auto pick = [&](auto nw) -> bool {
        if (decltype(nw)::value < need) return false;
        fn(nw);
        return true;
    };
    using std::integral_constant;
    bool ok;
    if constexpr (L == 2)
        ok = pick(integral_constant<int, 8>()) || pick(integral_constant<int, 12>()) ||
             pick(integral_constant<int, 16>());
    else if constexpr (L == 4)
        ok = pick(integral_constant<int, 12>()) || pick(integral_constant<int, 16>()) ||
             pick(integral_constant<int, 24>());
    else if constexpr (L == 8)
        ok = pick(integral_constant<int, 20>()) || pick(integral_constant<int, 28>()) ||
             pick(integral_constant<int, 40>());
    else
        ok = pick(integral_constant<int, 40>()) || pick(integral_constant<int, 56>()) ||
             pick(integral_constant<int, 92>());
    if (ok) return MDG_SUCCESS;
    set_error("no band reconstruction for %d slices at L=%d", S, L);
    return MDG_ERROR_RANGE;
}

// smallest word-count instantiation with at least `need` words; fn is
// called with std::integral_constant<int, NW>
template <int L, typename Fn>
inline int with_words(int need, Fn fn)
{
    auto pick = [&](auto nw) -> bool {
        if (decltype(nw)::value < need) return false;
        fn(nw);
        return true;
    };
    using std::integral_constant;
    if constexpr (L == 2) {
        if (pick(integral_constant<int, 12>()) || pick(integral_constant<int, 16>()) ||
            pick(integral_constant<int, 24>()) || pick(integral_constant<int, 32>()))
            return MDG_SUCCESS;
    } else if constexpr (L == 4) {
        if (pick(integral_constant<int, 16>()) || pick(integral_constant<int, 24>()) ||
            pick(integral_constant<int, 32>()) || pick(integral_constant<int, 48>()))
            return MDG_SUCCESS;
    } else if constexpr (L == 8) {
        if (pick(integral_constant<int, 16>()) || pick(integral_constant<int, 24>()) ||
            pick(integral_constant<int, 32>()) || pick(integral_constant<int, 48>()) ||
            pick(integral_constant<int, 64>()))
            return MDG_SUCCESS;
    } else {
        if (pick(integral_constant<int, 32>()) || pick(integral_constant<int, 48>()) ||
            pick(integral_constant<int, 64>()) || pick(integral_constant<int, 96>()))
            return MDG_SUCCESS;
    }
    set_error("no reconstruction kernel for %d diagonal-sum words at L=%d", need, L);
    return MDG_ERROR_RANGE;
}

template <int L, typename Fn>
inline int with_diag_words(int S, int w, Fn fn)
{
    return with_words<L>(diag_words_needed(S, w), fn);
}

// fn called with std::integral_constant<int, diag_group(w)>
template <typename Fn>
inline void with_group(int w, Fn fn)
{
    const int gs = diag_group(w);
    if (gs >= 3)
        fn(std::integral_constant<int, 3>());
    else if (gs == 2)
        fn(std::integral_constant<int, 2>());
    else
        fn(std::integral_constant<int, 1>());
}

}  // namespace mdg
