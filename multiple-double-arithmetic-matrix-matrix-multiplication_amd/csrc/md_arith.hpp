// md_arith.hpp -- multi-double arithmetic for gfx950 device code.
//
// Error-free transformations and the L-limb operations the GEMM path needs,
// with the operation order of the reference's scalar library (see the
// citations on each function), so that a kernel built from them produces the
// same bits as the reference's host loops:
//   two_sum / quick_two_sum   C++/double_double_functions.cu:18-24, :11-16
//   two_prod                  C++/double_double_functions.cu:150-184
//   dd add / mul / inc_d      C++/double_double_functions.cu:26-38, :197-206, :99-106
//   renorm ladder             C++/quad_double_functions.cu:14-152 (and the
//                             octo :15-426 / hexa :16-1362 instances)
//   fast_renorm, renorm_add1  quad :154-180, octo :427-465, hexa :1363-1425
//   add  (baileyAdd_fast)     quad :216-242, octo :515-576, hexa :1499-1688
//   mul  (baileyMul_fast)     quad :321-381, octo :723-1014, hexa :2044-3863
//
// Every translation unit including this is compiled with -ffp-contract=off:
// a contracted a*b+c would change the bits of every error term.
#pragma once
#include <hip/hip_runtime.h>

namespace mdg {

// two_prod variants: the reference's Dekker/Veltkamp product, or the
// hardware fma.  Both are exact (hence equal) unless a partial product
// underflows; fma is 2 instructions instead of 17.
enum TwoProd { TP_FMA = 0, TP_DEKKER = 1 };

__host__ __device__ __forceinline__ double quick_two_sum(double a, double b, double &e)
{
    double s = a + b;
    e = b - (s - a);
    return s;
}

__host__ __device__ __forceinline__ double two_sum(double a, double b, double &e)
{
    double s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
    return s;
}

__device__ __forceinline__ void veltkamp(double a, double &hi, double &lo)
{
    const double splitter = 134217729.0;           // 2^27 + 1
    const double thresh = 6.69692879491417e+299;   // 2^996
    if (a > thresh || a < -thresh) {
        a *= 3.7252902984619140625e-09;            // 2^-28
        double t = splitter * a;
        hi = t - (t - a);
        lo = a - hi;
        hi *= 268435456.0;                         // 2^28
        lo *= 268435456.0;
    } else {
        double t = splitter * a;
        hi = t - (t - a);
        lo = a - hi;
    }
}

template <int TP>
__device__ __forceinline__ double two_prod(double a, double b, double &e)
{
    double p = a * b;
    if constexpr (TP == TP_FMA) {
        e = __builtin_fma(a, b, -p);
    } else {
        double ah, al, bh, bl;
        veltkamp(a, ah, al);
        veltkamp(b, bh, bl);
        e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
    }
    return p;
}

// ---------------------------------------------------------------- L = 2 --
template <int TP>
__device__ __forceinline__ void dd_mul(const double *a, const double *b, double *c)
{
    double p1, p2;
    p1 = two_prod<TP>(a[0], b[0], p2);
    p2 += (a[0] * b[1] + a[1] * b[0]);
    double c1;
    double c0 = quick_two_sum(p1, p2, c1);
    c[0] = c0;
    c[1] = c1;
}

__device__ __forceinline__ void dd_add(const double *a, const double *b, double *c)
{
    double s1, s2, t1, t2;
    s1 = two_sum(a[0], b[0], s2);
    t1 = two_sum(a[1], b[1], t2);
    s2 += t1;
    s1 = quick_two_sum(s1, s2, s2);
    s2 += t2;
    double c1;
    double c0 = quick_two_sum(s1, s2, c1);
    c[0] = c0;
    c[1] = c1;
}

__device__ __forceinline__ void dd_inc_d(double *a, double y)
{
    double s1, s2;
    s1 = two_sum(a[0], y, s2);
    s2 += a[1];
    double a1;
    double a0 = quick_two_sum(s1, s2, a1);
    a[0] = a0;
    a[1] = a1;
}

// ------------------------------------------------------ renorm ladder --
// L+1 terms f[0..L] -> L limbs r[0..L-1] with the QD zero-skipping rule.
// The ladder keeps a write position that advances unless the running error
// is exactly zero.  For inputs whose intermediate errors are all non-zero
// (the common case) the position advances every step and the ladder is a
// straight chain of quick_two_sums; that chain is computed first and the
// general, position-tracking form runs only for lanes that hit a zero.
// The general form never indexes a register array with a run-time value:
// reads and writes of slot `ptr` are unrolled selects, so nothing spills.
template <int L>
__device__ __forceinline__ void renorm_slow(const double *f, double *r)
{
    double pr;
    int ptr;
    {
        const bool z = (f[1] == 0.0);
        double s = quick_two_sum(z ? f[0] : f[1], f[2], pr);
        r[0] = z ? s : f[0];
        r[1] = z ? 0.0 : s;
        ptr = z ? 0 : 1;
#pragma unroll
        for (int j = 2; j < L; ++j) r[j] = 0.0;
    }
#pragma unroll
    for (int i = 3; i <= L; ++i) {
        // before this step ptr <= i-2
        const bool z = (pr == 0.0);
        double cur = pr;
        if (z) {
            double v = r[0];
#pragma unroll
            for (int j = 1; j <= i - 2; ++j) v = (ptr == j) ? r[j] : v;
            cur = v;
        }
        ptr += z ? 0 : 1;
        double s = quick_two_sum(cur, f[i], pr);
#pragma unroll
        for (int j = 0; j <= i - 1; ++j) r[j] = (ptr == j) ? s : r[j];
    }
    {
        const bool z = (pr == 0.0);
        double cur = pr;
        if (z) {
            double v = r[0];
#pragma unroll
            for (int j = 1; j < L; ++j) v = (ptr == j) ? r[j] : v;
            cur = v;
        }
        ptr += z ? 0 : 1;
        if (ptr < L && cur != 0.0) {
#pragma unroll
            for (int j = 0; j < L; ++j) r[j] = (ptr == j) ? cur : r[j];
            ++ptr;
        }
#pragma unroll
        for (int j = 0; j < L; ++j) r[j] = (j >= ptr) ? 0.0 : r[j];
    }
}

template <int L>
__device__ __forceinline__ void renorm(const double *f, double *r)
{
    // straight chain: r0 = f0, then quick_two_sums down the terms
    double e[L + 1];
    double s[L];
    s[0] = f[0];
    e[1] = f[1];
    bool ok = (f[1] != 0.0);
#pragma unroll
    for (int i = 2; i <= L; ++i) {
        s[i - 1] = quick_two_sum(e[i - 1], f[i], e[i]);
        if (i <= L - 1) ok = ok && (e[i] != 0.0);
    }
    // the last error is dropped by the ladder in both forms
    if (__builtin_expect(ok, 1)) {
#pragma unroll
        for (int j = 0; j < L; ++j) r[j] = s[j];
    } else {
        renorm_slow<L>(f, r);
    }
}

template <int L>
__device__ __forceinline__ void fast_renorm(const double *x, double *r)
{
    double f[L + 1], pr;
    pr = quick_two_sum(x[L - 1], x[L], f[L]);
#pragma unroll
    for (int i = L - 2; i >= 1; --i) pr = quick_two_sum(x[i], pr, f[i + 1]);
    f[0] = quick_two_sum(x[0], pr, f[1]);
    renorm<L>(f, r);
}

template <int L>
__device__ __forceinline__ void renorm_add1(const double *x, double y, double *r)
{
    double f[L + 1], pr;
    pr = two_sum(x[L - 1], y, f[L]);
#pragma unroll
    for (int i = L - 2; i >= 1; --i) pr = two_sum(x[i], pr, f[i + 1]);
    f[0] = two_sum(x[0], pr, f[1]);
    renorm<L>(f, r);
}

// ------------------------------------------------------------ L-limb ops --
template <int L>
__device__ __forceinline__ void md_add(const double *a, const double *b, double *c)
{
    if constexpr (L == 2) {
        dd_add(a, b, c);
    } else {
        double f[L + 1], e;
        f[L] = 0.0;
#pragma unroll
        for (int n = L - 1; n >= 0; --n) {
            f[n] = two_sum(a[n], b[n], e);
#pragma unroll
            for (int m = n + 1; m < L; ++m) f[m] = two_sum(f[m], e, e);
            f[L] += e;
        }
        fast_renorm<L>(f, c);
    }
}

template <int L, int TP>
__device__ __forceinline__ void md_mul(const double *a, const double *b, double *c)
{
    if constexpr (L == 2) {
        dd_mul<TP>(a, b, c);
    } else {
        double f[L + 1], p, e;
        f[L] = a[1] * b[L - 1];
#pragma unroll
        for (int i = 2; i < L; ++i) f[L] += a[i] * b[L - i];
#pragma unroll
        for (int n = L - 1; n >= 0; --n) {
            f[n] = two_prod<TP>(a[0], b[n], e);
#pragma unroll
            for (int m = n + 1; m < L; ++m) f[m] = two_sum(f[m], e, e);
            f[L] += e;
#pragma unroll
            for (int i = 1; i <= n; ++i) {
                p = two_prod<TP>(a[i], b[n - i], e);
#pragma unroll
                for (int m = n + 1; m < L; ++m) f[m] = two_sum(f[m], e, e);
                f[L] += e;
                f[n] = two_sum(f[n], p, e);
#pragma unroll
                for (int m = n + 1; m < L; ++m) f[m] = two_sum(f[m], e, e);
                f[L] += e;
            }
        }
        fast_renorm<L>(f, c);
    }
}

// a += y for an L-limb a and a double y (inc_d)
template <int L>
__device__ __forceinline__ void md_inc_d(double *a, double y)
{
    if constexpr (L == 2) {
        dd_inc_d(a, y);
    } else {
        double r[L];
        renorm_add1<L>(a, y, r);
#pragma unroll
        for (int j = 0; j < L; ++j) a[j] = r[j];
    }
}

}  // namespace mdg
