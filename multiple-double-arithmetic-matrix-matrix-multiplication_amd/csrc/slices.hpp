// slices.hpp -- fixed-point slicing of multi-double numbers.
//
// An operand row (of A) or column (of B) with scale exponent E (every entry
// |x| < 2^E) is written as the integer N = floor(x * 2^(W-2-E)), |N| <
// 2^(W-2), and N is cut into S = W/w balanced base-2^w digits
//     N = sum_t d_t * 2^(w (S-1-t)),   d_t in [-2^(w-1), 2^(w-1) - 1],
// so that x ~= sum_t d_t * 2^(E + 2 - w (t+1)) with error below
// L * 2^(E + 2 - W).  Products of two digits are exact in every MFMA
// format used (INT8 x INT8 -> INT32, BF16 x BF16 -> FP32, FP64), and the
// slice-pair sums of one anti-diagonal g = t + u are exact integers, so the
// only roundings of the whole GEMM are the truncation of x to W bits and
// the final TwoSum renormalisation.
//
// N is built exactly in an NW-word two's-complement register array from the
// L limbs (each limb m * 2^e with a 53-bit signed m is added at bit
// e + W - 2 - E, truncating below bit 0), then digits are taken from the
// least significant end with a carry, which makes every digit balanced.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdg {

// words of the fixed-point register per limb count (W <= 32*NW - 1)
template <int L>
struct FixedWords {
    static constexpr int NW = (L == 2) ? 8 : (L == 4) ? 12 : (L == 8) ? 20 : 36;
};

// decompose a finite double: x = m * 2^e, |m| < 2^53 (m = 0 for zero)
__device__ __forceinline__ void limb_decomp(double x, long long &m, int &e)
{
    int ex = 0;
    double fr = frexp(x, &ex);            // |fr| in [0.5, 1) or 0
    m = (long long)ldexp(fr, 53);
    e = ex - 53;
}

// N += m * 2^p (p may be negative: floor), sign-extended over NW words
template <int NW>
__device__ __forceinline__ void fixed_add(uint32_t (&N)[NW], long long m, int p)
{
    if (m == 0) return;
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        const int q = 32 * j - p;      // bit of m landing on bit 0 of word j
        uint32_t chunk;
        if (q >= 63)
            chunk = (uint32_t)(m >> 63);
        else if (q >= 0)
            chunk = (uint32_t)(m >> q);
        else if (q > -32)
            chunk = (uint32_t)((uint64_t)m << (-q));
        else
            chunk = 0u;
        const uint64_t t = (uint64_t)N[j] + chunk + carry;
        N[j] = (uint32_t)t;
        carry = t >> 32;
    }
}

// bits [pos, pos + w) of N as an unsigned value (w <= 32)
template <int NW>
__device__ __forceinline__ uint32_t fixed_bits(const uint32_t (&N)[NW], int pos, int w)
{
    const int j = pos >> 5, off = pos & 31;
    uint64_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        lo = (i == j) ? N[i] : lo;
        hi = (i == j + 1) ? N[i] : hi;
    }
    if (j + 1 >= NW) hi = (N[NW - 1] >> 31) ? 0xFFFFFFFFull : 0ull;
    const uint64_t v = (lo | (hi << 32)) >> off;
    return (uint32_t)(v & ((w >= 32) ? 0xFFFFFFFFull : ((1ull << w) - 1)));
}

// Build N for one L-limb entry with shift = W - 2 - E.
template <int L, int NW>
__device__ __forceinline__ void to_fixed(const double *x, int shift, uint32_t (&N)[NW])
{
#pragma unroll
    for (int j = 0; j < NW; ++j) N[j] = 0u;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        long long m;
        int e;
        limb_decomp(x[l], m, e);
        fixed_add<NW>(N, m, e + shift);
    }
}

// Balanced digits, least significant first; emit(t, d) receives slice t.
// The words are streamed through a 64-bit bit buffer, so the cost is
// O(NW + S) per entry with no run-time register indexing except the one
// read of the top window.
template <int NW, typename Emit>
__device__ __forceinline__ void fixed_digits(const uint32_t (&N)[NW], int S, int w, Emit emit)
{
    const int half = 1 << (w - 1);
    const int full = 1 << w;
    const uint64_t mask = (1ull << w) - 1;
    int carry = 0;
    int t = S - 1;
    uint64_t buf = 0;
    int nb = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        if (t >= 1) {
            buf |= (uint64_t)N[j] << nb;   // nb < w <= 32 here
            nb += 32;
            while (nb >= w && t >= 1) {
                const int u = (int)(buf & mask) + carry;
                int d;
                if (u >= half) {
                    d = u - full;
                    carry = 1;
                } else {
                    d = u;
                    carry = 0;
                }
                emit(t, d);
                buf >>= w;
                nb -= w;
                --t;
            }
        }
    }
    // top digit: the signed w-bit window at w(S-1), plus the carry
    uint32_t raw = fixed_bits<NW>(N, w * (S - 1), w);
    int top = (int)raw;
    if (raw & (uint32_t)half) top -= full;
    emit(0, top + carry);
}

}  // namespace mdg
