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

// The exponent spread of a row or column: ceil(E - log2(mean |x0|)) with
// mean = sum0 / k, clamped to [0, 256].  With mean in [2^(f-1), 2^f) (f its
// frexp exponent), log2(mean) lies in [f-1, f), so the ceiling is exactly
// E - f + 1: taken from the two exponents, with no rounded logarithm, and
// unchanged when the row or column is scaled by a power of two.
__device__ __forceinline__ int exponent_spread(int e, double sum0, int k)
{
    int f;
    frexp(sum0 / (double)k, &f);
    return min(256, max(0, e - f + 1));
}

// words of the fixed-point register per limb count (W <= 32*NW - 1)
template <int L>
struct FixedWords {
    static constexpr int NW = (L == 2) ? 8 : (L == 4) ? 12 : (L == 8) ? 20 : 36;
};

// decompose a finite double: x = m * 2^e, |m| < 2^53 (m = 0 for zero), read
// from the bit fields (normal: the 53-bit significand with its implicit bit
// and e = biased exponent - 1075; subnormal: the raw 52 bits and e = -1074)
__device__ __forceinline__ void limb_decomp(double x, long long &m, int &e)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const int ex = (int)((u >> 52) & 0x7ff);
    const long long mant = (long long)((u & ((1ull << 52) - 1)) | ((ex != 0) ? (1ull << 52) : 0ull));
    e = ((ex != 0) ? ex : 1) - 1075;
    m = (u >> 63) ? -mant : mant;
}

// N += floor(m * 2^p), two's complement over NW words.  m 2^p = M 2^(32 j0)
// with b = p mod 32, j0 = floor(p / 32) and M = m 2^b: M's words w0, w1
// (bits 0..63) and w2 (bits 64..95, sign-extended), then its sign s in
// every word above.  Word j of N adds the word j - j0 of M (words below 0,
// the bits below N's bit 0, drop out: floor); straight-line selects and one
// add-with-carry chain, no branches.
template <int NW>
__device__ __forceinline__ void fixed_add(uint32_t (&N)[NW], long long m, int p)
{
    const int b = p & 31, j0 = p >> 5;
    const unsigned long long lo = (unsigned long long)m << b;
    const uint32_t s = (uint32_t)(m >> 63);
    const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)(lo >> 32);
    const uint32_t w2 = (uint32_t)((m >> 32) >> (32 - b));   // b = 0: the sign word
    unsigned int carry = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        const int d = j - j0;
        uint32_t a = (d >= 3) ? s : 0u;
        a = (d == 2) ? w2 : a;
        a = (d == 1) ? w1 : a;
        a = (d == 0) ? w0 : a;
        unsigned int co;
        N[j] = __builtin_addc(N[j], a, carry, &co);
        carry = co;
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

// w = 8 digits, the same as fixed_digits(N, S, 8, emit): with N' = N +
// sum_{i < S-1} 128 * 256^i, byte i of N' is d_i + 128 for the balanced
// digits d_i (i = 0 least significant) -- as an INT8, d_i = byte i xor 0x80
// -- and the top digit is N' >> 8 (S-1), a signed byte here (|N| < 2^(8S-2)).
// emit(t, d) receives slice t = S-1-i.  S must be uniform across the wave
// (the byte loop bounds are branches on it).
template <int NW, typename Emit>
__device__ __forceinline__ void fixed_digits8(uint32_t (&N)[NW], int S, Emit emit)
{
    const int nb = S - 1;   // bytes that carry the +128 offset
    unsigned int carry = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        const int k = nb - 4 * j;   // offset bytes in word j
        const uint32_t add = (k >= 4) ? 0x80808080u : ((k > 0) ? (0x80808080u >> (8 * (4 - k))) : 0u);
        unsigned int co;
        N[j] = __builtin_addc(N[j], add, carry, &co);
        carry = co;
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        if (4 * j > nb) break;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 4 * j + q;
            const int byte = (int)((N[j] >> (8 * q)) & 0xffu);
            if (i < nb)
                emit(nb - i, byte - 128);
            else if (i == nb)
                emit(0, (byte ^ 0x80) - 128);   // the signed byte
        }
    }
}

}  // namespace mdg
