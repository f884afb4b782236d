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

// ------------------------------------------- operations off the GEMM path --
// sub, sqr, mul by a double, div, sqrt with the reference's operation order:
//   dd:  ddf_sub :61-73, ddf_two_sqr :186-195, ddf_sqr :208-216,
//        ddf_mul_d_dd :218-228, ddf_div :230-248, ddf_sqrt :250-279
//   qd/od/hd: X_sub (copy, minus, inc) quad :299-309, octo :684-704,
//        hexa :1957-2005; X_mul_X_d (baileyMul_fast<L,1,L>) quad :442-468,
//        octo :1305-1365, hexa :5676-5858; X_div quad :485-520, octo
//        :1390-1505, hexa :5903-6296; X_sqrt quad :524-537, octo :1507-1536,
//        hexa :6298-6361.  The qd/od/hd sqr bodies are the mul bodies with
//        b = a (quad :383-440 against :321-381), so md_sqr is md_mul(a, a).

__device__ __forceinline__ double two_diff(double a, double b, double &e)
{
    double s = a - b;
    double bb = s - a;
    e = (a - (s - bb)) - (b + bb);
    return s;
}

__device__ __forceinline__ void dd_sub(const double *a, const double *b, double *c)
{
    double s1, s2, t1, t2;
    s1 = two_diff(a[0], b[0], s2);
    t1 = two_diff(a[1], b[1], t2);
    s2 += t1;
    s1 = quick_two_sum(s1, s2, s2);
    s2 += t2;
    double c1;
    double c0 = quick_two_sum(s1, s2, c1);
    c[0] = c0;
    c[1] = c1;
}

template <int TP>
__device__ __forceinline__ void dd_sqr(const double *a, double *c)
{
    double p2;
    const double p1 = a[0] * a[0];
    if constexpr (TP == TP_FMA) {
        p2 = __builtin_fma(a[0], a[0], -p1);
    } else {
        double hi, lo;
        veltkamp(a[0], hi, lo);
        p2 = ((hi * hi - p1) + 2.0 * hi * lo) + lo * lo;
    }
    p2 += 2.0 * a[0] * a[1];
    p2 += a[1] * a[1];
    double c1;
    double c0 = quick_two_sum(p1, p2, c1);
    c[0] = c0;
    c[1] = c1;
}

// c = a * d for a double d
template <int L, int TP>
__device__ __forceinline__ void md_mul_d(const double *a, double d, double *c)
{
    if constexpr (L == 2) {
        double p2;
        const double p1 = two_prod<TP>(d, a[0], p2);
        p2 += (a[1] * d);
        double c1;
        double c0 = quick_two_sum(p1, p2, c1);
        c[0] = c0;
        c[1] = c1;
    } else {
        double f[L + 1], e;
        f[L] = 0.0;
        f[L - 1] = two_prod<TP>(a[L - 1], d, e);
        f[L] += e;
#pragma unroll
        for (int i = L - 2; i >= 0; --i) {
            f[i] = two_prod<TP>(a[i], d, e);
#pragma unroll
            for (int j = i + 1; j < L; ++j) f[j] = two_sum(f[j], e, e);
            f[L] += e;
        }
        fast_renorm<L>(f, c);
    }
}

template <int L>
__device__ __forceinline__ void md_sub(const double *a, const double *b, double *c)
{
    if constexpr (L == 2) {
        dd_sub(a, b, c);
    } else {
        double t[L];
#pragma unroll
        for (int i = 0; i < L; ++i) t[i] = -b[i];
        md_add<L>(t, a, c);
    }
}

template <int L, int TP>
__device__ __forceinline__ void md_sqr(const double *a, double *c)
{
    if constexpr (L == 2)
        dd_sqr<TP>(a, c);
    else
        md_mul<L, TP>(a, a, c);
}

// quotient digits q_i = r_0 / b_0, r -= q_i b: three for dd (the last two
// folded in by quick_two_sum and inc_d), L + 1 through fast_renorm otherwise
template <int L, int TP>
__device__ __forceinline__ void md_div(const double *a, const double *b, double *c)
{
    double acc[L], r[L], t[L];
    if constexpr (L == 2) {
        const double q1 = a[0] / b[0];
        md_mul_d<2, TP>(b, q1, acc);
        dd_sub(a, acc, r);
        const double q2 = r[0] / b[0];
        md_mul_d<2, TP>(b, q2, acc);
        dd_sub(r, acc, t);
        const double q3 = t[0] / b[0];
        double c1;
        double c0 = quick_two_sum(q1, q2, c1);
        c[0] = c0;
        c[1] = c1;
        dd_inc_d(c, q3);
    } else {
        double q[L + 1];
        q[0] = a[0] / b[0];
        md_mul_d<L, TP>(b, q[0], acc);
        md_sub<L>(a, acc, r);
#pragma unroll
        for (int i = 1; i < L; ++i) {
            q[i] = r[0] / b[0];
            md_mul_d<L, TP>(b, q[i], acc);
            md_sub<L>(r, acc, t);
#pragma unroll
            for (int j = 0; j < L; ++j) r[j] = t[j];
        }
        q[L] = r[0] / b[0];
        fast_renorm<L>(q, c);
    }
}

// dd: Karp's step from 1/sqrt(a_0), 0 for 0 and -1 for a negative leading
// limb (the reference's convention); L >= 4: the root s of the leading L/2
// limbs, then one Heron step (s^2 + a) / s * 0.5 at full length
template <int L, int TP>
__device__ __forceinline__ void md_sqrt(const double *a, double *b)
{
    if constexpr (L == 2) {
        if (a[0] == 0.0 && a[1] == 0.0) {
            b[0] = 0.0;
            b[1] = 0.0;
        } else if (a[0] < 0.0) {
            b[0] = -1.0;
            b[1] = 0.0;
        } else {
            const double x = 1.0 / sqrt(a[0]);
            const double ax = a[0] * x;
            const double s[2] = {ax, 0.0};
            double sq[2], y[2];
            dd_sqr<TP>(s, sq);
            dd_sub(a, sq, y);
            double b1;
            const double b0 = two_sum(ax, y[0] * x * 0.5, b1);
            b[0] = b0;
            b[1] = b1;
        }
    } else {
        double s[L], z[L], t[L];
        md_sqrt<L / 2, TP>(a, s);
#pragma unroll
        for (int i = L / 2; i < L; ++i) s[i] = 0.0;
        md_mul<L, TP>(s, s, z);
        md_add<L>(z, a, t);
        md_div<L, TP>(t, s, z);
#pragma unroll
        for (int i = 0; i < L; ++i) b[i] = z[i] * 0.5;
    }
}

}  // namespace mdg
