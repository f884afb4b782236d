/*
 * md_oracle.c -- CPU oracle for the multiple-double GEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker (or as the
 * timed CPU baseline), never as the thing measured on the GPU.
 *
 * It restates, in plain C, the arithmetic of the reference repository
 * (hchen221/multiple-double-arithmetic-matrix-matrix-multiplication, C++/):
 *
 *   L0  error-free transformations
 *       two_sum            C++/double_double_functions.cu:18-24
 *       quick_two_sum      C++/double_double_functions.cu:11-16
 *       Veltkamp split     C++/double_double_functions.cu:150-172
 *       Dekker two_prod    C++/double_double_functions.cu:174-184
 *   L1  double-double      ddf_add :26-38, ddf_mul :197-206, ddf_inc_d :99-106
 *       L-limb (4/8/16)    renorm ladder     quad :14-152, octo :15-426,
 *                                            hexa :16-1362
 *                          fast_renorm       quad :154-166, octo :427-445,
 *                                            hexa :1363-1393
 *                          renorm_add1       quad :168-180, octo :446-465,
 *                                            hexa :1394-1425
 *                          add (baileyAdd_fast<L,L,L>)  quad :216-242,
 *                                            octo :515-576, hexa :1499-1688
 *                          mul (baileyMul_fast<L,L,L>)  quad :321-381,
 *                                            octo :723-1014, hexa :2044-3863
 *                          inc_d             quad :269-275, octo :637-646,
 *                                            hexa :1871-1890
 *                          sub, sqr, mul by a double, div, sqrt, abs, minus,
 *                          mul_pwr2          double :47-73,186-279,
 *                                            quad :194-212,290-537,
 *                                            octo :669-722,1016-1536,
 *                                            hexa :1937-2043,3865-6361
 *   L3  host GEMM          matmul{dd,qd,od,hd}f  C++/TCcuda.cu:320-372
 *       squeeze            pllsqueeze_{2,4,8,16}_pp_old  C++/TCcuda.cu:876-906
 *       limb convolution   bigB2 + WMMA matmul  C++/TCcuda.cu:79-95,
 *                          C++/TCtest.cu:21-46; in plain doubles in the
 *                          reference's loop order, matmulhost
 *                          C++/TCcuda.cu:304-317 (terms pinned by exact
 *                          small-integer products, the order by hand-built
 *                          rounding cases; tests/test_oracle.py)
 *       renormhost         renormhost / renormhostbig  C++/TCcuda.cu:374-398
 *                          (pinned by closed forms: exact sums give
 *                          x_k + x_{k+1}, a lost tail part gives 0)
 *   L2  splitter           bitform, double_rep, split4, split8, balance,
 *                          split4pd/split8pd/mixsplit2  C++/floatybits.cu:54-342
 *       generator          random_pd  C++/floatybits.cu:73-120 (own PRNG)
 *
 * The reference's generated code for each L is restated once, as loops over
 * the limb index with L as a parameter; the operation order of every loop is
 * the order the reference's unrolled code performs.
 *
 * It also holds an exact checker: a big-integer fixed-point dot product that
 * represents every partial product exactly (no rounding anywhere), used to
 * measure normwise error of any GEMM result.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off, OpenMP).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXL 16

/* ------------------------------------------------------------------ */
/* L0: error-free transformations                                      */
/* ------------------------------------------------------------------ */

/* Dekker's FastTwoSum: s + e == a + b exactly when |a| >= |b|.
   double_double_functions.cu:11-16 */
static inline double qts(double a, double b, double *e)
{
    double s = a + b;
    *e = b - (s - a);
    return s;
}

/* Knuth's TwoSum: s + e == a + b exactly.  double_double_functions.cu:18-24 */
static inline double ts(double a, double b, double *e)
{
    double s = a + b;
    double bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}

/* Veltkamp split with a 2^996 overflow guard.
   double_double_functions.cu:150-172 */
static inline void vsplit(double a, double *hi, double *lo)
{
    const double splitter = 134217729.0;            /* 2^27 + 1 */
    const double thresh = 6.69692879491417e+299;    /* 2^996 */
    double t;
    if (a > thresh || a < -thresh) {
        a *= 3.7252902984619140625e-09;             /* 2^-28 */
        t = splitter * a;
        *hi = t - (t - a);
        *lo = a - *hi;
        *hi *= 268435456.0;                         /* 2^28 */
        *lo *= 268435456.0;
    } else {
        t = splitter * a;
        *hi = t - (t - a);
        *lo = a - *hi;
    }
}

/* Dekker's TwoProduct.  double_double_functions.cu:174-184 */
static inline double tp_dekker(double a, double b, double *e)
{
    double ah, al, bh, bl;
    double p = a * b;
    vsplit(a, &ah, &al);
    vsplit(b, &bh, &bl);
    *e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
    return p;
}

/* TwoProduct through a fused multiply-add.  Equal to tp_dekker whenever no
   partial product underflows or overflows (both are then exact). */
static inline double tp_fma(double a, double b, double *e)
{
    double p = a * b;
    *e = fma(a, b, -p);
    return p;
}

static inline double tp(double a, double b, double *e, int mode)
{
    return mode ? tp_fma(a, b, e) : tp_dekker(a, b, e);
}

/* ------------------------------------------------------------------ */
/* L1: double-double (QD algorithms)                                   */
/* ------------------------------------------------------------------ */

/* ddf_add, double_double_functions.cu:26-38 */
static inline void dd_add(const double *a, const double *b, double *c)
{
    double s1, s2, t1, t2;
    s1 = ts(a[0], b[0], &s2);
    t1 = ts(a[1], b[1], &t2);
    s2 += t1;
    s1 = qts(s1, s2, &s2);
    s2 += t2;
    double c0, c1;
    c0 = qts(s1, s2, &c1);
    c[0] = c0;
    c[1] = c1;
}

/* ddf_mul, double_double_functions.cu:197-206 */
static inline void dd_mul(const double *a, const double *b, double *c, int mode)
{
    double p1, p2;
    p1 = tp(a[0], b[0], &p2, mode);
    p2 += (a[0] * b[1] + a[1] * b[0]);
    double c0, c1;
    c0 = qts(p1, p2, &c1);
    c[0] = c0;
    c[1] = c1;
}

/* ddf_inc_d, double_double_functions.cu:99-106 */
static inline void dd_inc_d(double *a, double b)
{
    double s1, s2;
    s1 = ts(a[0], b, &s2);
    s2 += a[1];
    double a0, a1;
    a0 = qts(s1, s2, &a1);
    a[0] = a0;
    a[1] = a1;
}

/* ------------------------------------------------------------------ */
/* L1: L-limb operations, L in {4, 8, 16}                              */
/* ------------------------------------------------------------------ */

/* QD renormalisation with zero skipping: L+1 terms f[0..L] -> L limbs.
   quad_double_functions.cu:14-152 (L=4), octo_double_functions.cu:15-426
   (L=8), hexa_double_functions.cu:16-1362 (L=16).  The reference spells out
   every value of the write position `ptr` as an if-chain; here it indexes. */
static inline void renorm_ladder(int L, const double *f, double *r)
{
    double pr;
    int ptr;
    if (f[1] == 0.0) {
        pr = f[0];
        ptr = 0;
        r[0] = qts(pr, f[2], &pr);
    } else {
        r[0] = f[0];
        pr = f[1];
        ptr = 1;
        r[1] = qts(pr, f[2], &pr);
    }
    for (int i = 3; i <= L; ++i) {
        if (pr == 0.0)
            pr = r[ptr];
        else
            ++ptr;
        r[ptr] = qts(pr, f[i], &pr);
    }
    if (pr == 0.0)
        pr = r[ptr];
    else
        ++ptr;
    if (ptr < L && pr != 0.0) {
        r[ptr] = pr;
        ++ptr;
    }
    for (int i = ptr; i < L; ++i)
        r[i] = 0.0;
}

/* Bottom-up QuickTwoSum sweep then the ladder.
   quad :154-166, octo :427-445, hexa :1363-1393 */
static inline void fast_renorm(int L, const double *x, double *r)
{
    double f[MAXL + 1], pr;
    pr = qts(x[L - 1], x[L], &f[L]);
    for (int i = L - 2; i >= 1; --i)
        pr = qts(x[i], pr, &f[i + 1]);
    f[0] = qts(x[0], pr, &f[1]);
    renorm_ladder(L, f, r);
}

/* Add one double to an L-limb number.  quad :168-180, octo :446-465,
   hexa :1394-1425; inc_d is this with r == x (quad :269-275). */
static inline void renorm_add1(int L, const double *x, double y, double *r)
{
    double f[MAXL + 1], pr;
    pr = ts(x[L - 1], y, &f[L]);
    for (int i = L - 2; i >= 1; --i)
        pr = ts(x[i], pr, &f[i + 1]);
    f[0] = ts(x[0], pr, &f[1]);
    renorm_ladder(L, f, r);
}

/* CAMPARY baileyAdd_fast<L,L,L>: limb-wise TwoSums from the least
   significant limb up, each error rippled down through the lower partial
   sums, the last term collecting plain sums.  quad :216-242 */
static inline void md_add(int L, const double *a, const double *b, double *c)
{
    double f[MAXL + 1], e;
    f[L] = 0.0;
    for (int n = L - 1; n >= 0; --n) {
        f[n] = ts(a[n], b[n], &e);
        for (int m = n + 1; m < L; ++m)
            f[m] = ts(f[m], e, &e);
        f[L] += e;
    }
    fast_renorm(L, f, c);
}

/* CAMPARY baileyMul_fast<L,L,L>: the last term is the plain sum of the
   products a_i*b_{L-i}; then for every anti-diagonal n = L-1 .. 0 the exact
   products a_i*b_{n-i} (i ascending) are folded in with TwoSums, each
   rounding error rippled through f[n+1..L-1] and collected in f[L].
   quad :321-381, octo :723-1014, hexa :2044-3863 */
static inline void md_mul(int L, const double *a, const double *b, double *c,
                          int mode)
{
    double f[MAXL + 1], p, e;
    f[L] = a[1] * b[L - 1];
    for (int i = 2; i < L; ++i)
        f[L] += a[i] * b[L - i];
    for (int n = L - 1; n >= 0; --n) {
        f[n] = tp(a[0], b[n], &e, mode);
        for (int m = n + 1; m < L; ++m)
            f[m] = ts(f[m], e, &e);
        f[L] += e;
        for (int i = 1; i <= n; ++i) {
            p = tp(a[i], b[n - i], &e, mode);
            for (int m = n + 1; m < L; ++m)
                f[m] = ts(f[m], e, &e);
            f[L] += e;
            f[n] = ts(f[n], p, &e);
            for (int m = n + 1; m < L; ++m)
                f[m] = ts(f[m], e, &e);
            f[L] += e;
        }
    }
    fast_renorm(L, f, c);
}

/* Dispatch on L: L == 2 uses the QD double-double algorithms. */
static inline void any_add(int L, const double *a, const double *b, double *c)
{
    if (L == 2)
        dd_add(a, b, c);
    else
        md_add(L, a, b, c);
}

static inline void any_mul(int L, const double *a, const double *b, double *c,
                           int mode)
{
    if (L == 2)
        dd_mul(a, b, c, mode);
    else
        md_mul(L, a, b, c, mode);
}

static inline void any_inc_d(int L, double *a, double y)
{
    if (L == 2) {
        dd_inc_d(a, y);
    } else {
        double r[MAXL];
        renorm_add1(L, a, y, r);
        memcpy(a, r, sizeof(double) * (size_t)L);
    }
}

/* ------------------------------------------------------------------ */
/* L1: the scalar operations off the GEMM path: sub, sqr, mul by a     */
/* double, div, sqrt (plus the trivial abs, negation, power-of-two     */
/* scaling).  The reference spells them out per L; the qd/od/hd sqr    */
/* bodies are the mul bodies with b replaced by a, and the qd/od/hd    */
/* inc bodies are the add bodies, so both are restated as those.       */
/* ------------------------------------------------------------------ */

/* two_diff, double_double_functions.cu:47-53: s + e == a - b exactly */
static inline double tdiff(double a, double b, double *e)
{
    double s = a - b;
    double bb = s - a;
    *e = (a - (s - bb)) - (b + bb);
    return s;
}

/* ddf_sub :61-73 (ddf_dec :108-118 is the same with c == a) */
static inline void dd_sub(const double *a, const double *b, double *c)
{
    double s1, s2, t1, t2;
    s1 = tdiff(a[0], b[0], &s2);
    t1 = tdiff(a[1], b[1], &t2);
    s2 += t1;
    s1 = qts(s1, s2, &s2);
    s2 += t2;
    double c0, c1;
    c0 = qts(s1, s2, &c1);
    c[0] = c0;
    c[1] = c1;
}

/* ddf_two_sqr :186-195 (Dekker), or through fma */
static inline double two_sqr(double a, double *e, int mode)
{
    double q = a * a;
    if (mode) {
        *e = fma(a, a, -q);
    } else {
        double hi, lo;
        vsplit(a, &hi, &lo);
        *e = ((hi * hi - q) + 2.0 * hi * lo) + lo * lo;
    }
    return q;
}

/* ddf_sqr :208-216 */
static inline void dd_sqr(const double *a, double *c, int mode)
{
    double p1, p2;
    p1 = two_sqr(a[0], &p2, mode);
    p2 += 2.0 * a[0] * a[1];
    p2 += a[1] * a[1];
    double c0, c1;
    c0 = qts(p1, p2, &c1);
    c[0] = c0;
    c[1] = c1;
}

/* ddf_mul_d_dd :218-228, c = d * b for a double d */
static inline void dd_mul_d(double d, const double *b, double *c, int mode)
{
    double p1, p2;
    p1 = tp(d, b[0], &p2, mode);
    p2 += (b[1] * d);
    double c0, c1;
    c0 = qts(p1, p2, &c1);
    c[0] = c0;
    c[1] = c1;
}

/* ddf_div :230-248: three quotient digits */
static inline void dd_div(const double *a, const double *b, double *c, int mode)
{
    double q1, q2, q3, acc[2], r[2], t[2];
    q1 = a[0] / b[0];
    dd_mul_d(q1, b, acc, mode);
    dd_sub(a, acc, r);
    q2 = r[0] / b[0];
    dd_mul_d(q2, b, acc, mode);
    dd_sub(r, acc, t);
    q3 = t[0] / b[0];
    double c0, c1;
    c0 = qts(q1, q2, &c1);
    c[0] = c0;
    c[1] = c1;
    dd_inc_d(c, q3);
}

/* ddf_sqrt :250-279: Karp's step from 1/sqrt(a_hi); zero gives zero and a
   negative leading limb gives -1 (the reference's convention) */
static inline void dd_sqrt(const double *a, double *b, int mode)
{
    if (a[0] == 0.0 && a[1] == 0.0) {
        b[0] = 0.0;
        b[1] = 0.0;
    } else if (a[0] < 0.0) {
        b[0] = -1.0;
        b[1] = 0.0;
    } else {
        const double x = 1.0 / sqrt(a[0]);
        const double ax = a[0] * x;
        double s[2] = {ax, 0.0}, sq[2], y[2];
        dd_sqr(s, sq, mode);
        dd_sub(a, sq, y);
        double b0, b1;
        b0 = ts(ax, y[0] * x * 0.5, &b1);
        b[0] = b0;
        b[1] = b1;
    }
}

/* qdf_sub :299-309, odf_sub :684-704, hdf_sub :1957-2005:
   c = -b, then c += a (inc, whose body is add's) */
static inline void md_sub(int L, const double *a, const double *b, double *c)
{
    double t[MAXL];
    for (int i = 0; i < L; ++i)
        t[i] = -b[i];
    md_add(L, t, a, c);
}

/* CAMPARY baileyMul_fast<L,1,L>: qdf_mul_qd_d :442-468, odf_mul_od_d
   :1305-1365, hdf_mul_hd_d :5676-5858.  Limb products from the least
   significant up, each error rippled down the lower partial sums. */
static inline void md_mul_d(int L, const double *a, double d, double *c, int mode)
{
    double f[MAXL + 1], e;
    f[L] = 0.0;
    f[L - 1] = tp(a[L - 1], d, &e, mode);
    f[L] += e;
    for (int i = L - 2; i >= 0; --i) {
        f[i] = tp(a[i], d, &e, mode);
        for (int j = i + 1; j < L; ++j)
            f[j] = ts(f[j], e, &e);
        f[L] += e;
    }
    fast_renorm(L, f, c);
}

/* qdf_div :485-520, odf_div :1390-1505, hdf_div :5903-6296: L+1 quotient
   digits q_i = r_0 / b_0 with r -= q_i b, then fast_renorm of the digits */
static inline void md_div(int L, const double *a, const double *b, double *c, int mode)
{
    double q[MAXL + 1], acc[MAXL], r[MAXL], t[MAXL];
    q[0] = a[0] / b[0];
    md_mul_d(L, b, q[0], acc, mode);
    md_sub(L, a, acc, r);
    for (int i = 1; i < L; ++i) {
        q[i] = r[0] / b[0];
        md_mul_d(L, b, q[i], acc, mode);
        md_sub(L, r, acc, t);
        memcpy(r, t, sizeof(double) * (size_t)L);
    }
    q[L] = r[0] / b[0];
    fast_renorm(L, q, c);
}

static void any_sqrt(int L, const double *a, double *b, int mode);

/* qdf_sqrt :524-537, odf_sqrt :1507-1536, hdf_sqrt :6298-6361: the root of
   the leading L/2 limbs at half precision, then one Heron step
   b = (s^2 + a) / s * 0.5 with s that root padded with zero limbs */
static inline void md_sqrt(int L, const double *a, double *b, int mode)
{
    double s[MAXL], z[MAXL], t[MAXL];
    const int h = L / 2;
    any_sqrt(h, a, s, mode);
    for (int i = h; i < L; ++i)
        s[i] = 0.0;
    md_mul(L, s, s, z, mode);
    md_add(L, z, a, t);
    md_div(L, t, s, z, mode);
    for (int i = 0; i < L; ++i)
        b[i] = z[i] * 0.5;
}

static void any_sqrt(int L, const double *a, double *b, int mode)
{
    if (L == 2)
        dd_sqrt(a, b, mode);
    else
        md_sqrt(L, a, b, mode);
}

/* Operation codes, equal to the product library's mdg_op values. */
enum {
    OP_ADD = 0, OP_SUB = 1, OP_MUL = 2, OP_DIV = 3, OP_SQR = 4, OP_SQRT = 5,
    OP_MUL_D = 6, OP_INC_D = 7, OP_MUL_PWR2 = 8, OP_ABS = 9, OP_NEG = 10
};

/* One entry of an elementwise operation: a and c L-limb, b L-limb for the
   binary ops, b[0] a double for the _D/_PWR2 ops, unused otherwise. */
static int one_op(int op, int L, const double *a, const double *b, double *c, int mode)
{
    double t[MAXL];
    switch (op) {
    case OP_ADD: any_add(L, a, b, c); break;
    case OP_SUB:
        if (L == 2) dd_sub(a, b, c);
        else md_sub(L, a, b, c);
        break;
    case OP_MUL: any_mul(L, a, b, c, mode); break;
    case OP_DIV:
        if (L == 2) dd_div(a, b, c, mode);
        else md_div(L, a, b, c, mode);
        break;
    case OP_SQR:
        if (L == 2) dd_sqr(a, c, mode);
        else md_mul(L, a, a, c, mode);
        break;
    case OP_SQRT: any_sqrt(L, a, c, mode); break;
    case OP_MUL_D:
        if (L == 2) dd_mul_d(b[0], a, c, mode);
        else md_mul_d(L, a, b[0], c, mode);
        break;
    case OP_INC_D:
        memcpy(t, a, sizeof(double) * (size_t)L);
        any_inc_d(L, t, b[0]);
        memcpy(c, t, sizeof(double) * (size_t)L);
        break;
    case OP_MUL_PWR2:   /* qdf_mul_pwr2 :311-319: limb-wise, exact */
        for (int i = 0; i < L; ++i) c[i] = a[i] * b[0];
        break;
    case OP_ABS:        /* qdf_abs :194-212: sign of the leading limb */
        for (int i = 0; i < L; ++i) c[i] = a[0] < 0.0 ? -a[i] : a[i];
        break;
    case OP_NEG:        /* qdf_minus :290-297 */
        for (int i = 0; i < L; ++i) c[i] = -a[i];
        break;
    default: return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Exported per-operation entry points (vectorised over `count`).      */
/* ------------------------------------------------------------------ */

static int valid_L(int L)
{
    return L == 2 || L == 4 || L == 8 || L == 16;
}

int orc_add(int L, long count, const double *a, const double *b, double *c)
{
    if (!valid_L(L)) return 1;
    for (long i = 0; i < count; ++i)
        any_add(L, a + i * L, b + i * L, c + i * L);
    return 0;
}

int orc_mul(int L, long count, const double *a, const double *b, double *c,
            int mode)
{
    if (!valid_L(L)) return 1;
    for (long i = 0; i < count; ++i)
        any_mul(L, a + i * L, b + i * L, c + i * L, mode);
    return 0;
}

/* inc_d: a[i] += y[i] (in place on a copy written to c). */
int orc_inc_d(int L, long count, const double *a, const double *y, double *c)
{
    if (!valid_L(L)) return 1;
    for (long i = 0; i < count; ++i) {
        double t[MAXL];
        memcpy(t, a + i * L, sizeof(double) * (size_t)L);
        any_inc_d(L, t, y[i]);
        memcpy(c + i * L, t, sizeof(double) * (size_t)L);
    }
    return 0;
}

/* fast_renorm / renorm ladder on L+1 inputs (L in {4,8,16}). */
int orc_fast_renorm(int L, long count, const double *x, double *r)
{
    if (L != 4 && L != 8 && L != 16) return 1;
    for (long i = 0; i < count; ++i)
        fast_renorm(L, x + i * (L + 1), r + i * L);
    return 0;
}

int orc_renorm(int L, long count, const double *f, double *r)
{
    if (L != 4 && L != 8 && L != 16) return 1;
    for (long i = 0; i < count; ++i)
        renorm_ladder(L, f + i * (L + 1), r + i * L);
    return 0;
}

/* Elementwise operation `op` over `count` entries (operand strides: L for
   a, c and the binary ops' b, 1 for the double operand of the _D/_PWR2
   ops). */
int orc_op(int op, int L, long count, const double *a, const double *b, double *c, int mode)
{
    if (!valid_L(L) || op < OP_ADD || op > OP_NEG) return 1;
    const int bstride = (op == OP_MUL_D || op == OP_INC_D || op == OP_MUL_PWR2) ? 1
                        : (op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV) ? L : 0;
    for (long i = 0; i < count; ++i)
        if (one_op(op, L, a + i * L, b ? b + i * bstride : NULL, c + i * L, mode)) return 1;
    return 0;
}

int orc_two_prod(long count, const double *a, const double *b, double *p,
                 double *e, int mode)
{
    for (long i = 0; i < count; ++i)
        p[i] = tp(a[i], b[i], &e[i], mode);
    return 0;
}

/* ------------------------------------------------------------------ */
/* L3: host GEMM, the reference's CPU path.  C++/TCcuda.cu:320-372      */
/*     C[r][c] = fold_k add(acc, mul(A[r][k], B[k][c])), acc starting   */
/*     at zero, k ascending; interleaved limbs, lda/ldb/ldc in entries. */
/* ------------------------------------------------------------------ */

#define DEFINE_GEMM_ROW(LL)                                                   \
    static void gemm_rows_##LL(int r0, int r1, int n, int k, const double *A, \
                               long lda, const double *B, long ldb,           \
                               double *C, long ldc, int mode)                 \
    {                                                                         \
        for (int r = r0; r < r1; ++r)                                         \
            for (int c = 0; c < n; ++c) {                                     \
                double acc[LL], prod[LL];                                     \
                for (int i = 0; i < LL; ++i) acc[i] = 0.0;                    \
                for (int kk = 0; kk < k; ++kk) {                              \
                    any_mul(LL, A + ((long)r * lda + kk) * LL,                \
                            B + ((long)kk * ldb + c) * LL, prod, mode);       \
                    any_add(LL, acc, prod, acc);                              \
                }                                                             \
                memcpy(C + ((long)r * ldc + c) * LL, acc, sizeof(acc));       \
            }                                                                 \
    }
DEFINE_GEMM_ROW(2)
DEFINE_GEMM_ROW(4)
DEFINE_GEMM_ROW(8)
DEFINE_GEMM_ROW(16)

/* Rows [r0, r1) of C; nthreads > 1 splits the rows over OpenMP threads. */
int orc_gemm_rows(int L, int r0, int r1, int n, int k, const double *A,
                  long lda, const double *B, long ldb, double *C, long ldc,
                  int mode, int nthreads)
{
    if (!valid_L(L) || r0 < 0 || r1 < r0 || n < 0 || k < 0) return 1;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int r = r0; r < r1; ++r) {
        switch (L) {
        case 2: gemm_rows_2(r, r + 1, n, k, A, lda, B, ldb, C, ldc, mode); break;
        case 4: gemm_rows_4(r, r + 1, n, k, A, lda, B, ldb, C, ldc, mode); break;
        case 8: gemm_rows_8(r, r + 1, n, k, A, lda, B, ldb, C, ldc, mode); break;
        default: gemm_rows_16(r, r + 1, n, k, A, lda, B, ldb, C, ldc, mode); break;
        }
    }
    return 0;
}

int orc_gemm(int L, int m, int n, int k, const double *A, long lda,
             const double *B, long ldb, double *C, long ldc, int mode,
             int nthreads)
{
    return orc_gemm_rows(L, 0, m, n, k, A, lda, B, ldb, C, ldc, mode, nthreads);
}

/* ------------------------------------------------------------------ */
/* Squeeze: y = fold_j inc_d(y, x[j]) over pp parts, y starting at 0.  */
/* C++/TCcuda.cu:876-906 (pllsqueeze_{2,4,8,16}_pp_old)                */
/* ------------------------------------------------------------------ */
int orc_squeeze(int L, long nent, int pp, const double *x, double *y)
{
    if (!valid_L(L) || pp < 1) return 1;
    for (long i = 0; i < nent; ++i) {
        double acc[MAXL];
        for (int l = 0; l < L; ++l) acc[l] = 0.0;
        for (int j = 0; j < pp; ++j)
            any_inc_d(L, acc, x[i * pp + j]);
        memcpy(y + i * L, acc, sizeof(double) * (size_t)L);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Limb convolution GEMM: the product Aq * bigB2(Bq) of the reference's */
/* split path (C++/TCcuda.cu:79-95 builds the block-Toeplitz operand,   */
/* C++/TCtest.cu:21-46 multiplies):                                     */
/*   C1q[r][c][i] = sum_k sum_{j<=i} Aq[r][k][j] * Bq[k][c][i-j]        */
/* Plain double sums (k outer, j inner); exact whenever every partial   */
/* sum is representable, the domain the reference's path relies on.    */
/* The same loop, in the same order, is the reference's host function  */
/* matmulhost (C++/TCcuda.cu:304-317, square n, p parts).              */
/* ------------------------------------------------------------------ */
int orc_conv_gemm(int m, int n, int k, int pp, const double *Aq,
                  const double *Bq, double *C1q)
{
    if (pp < 1) return 1;
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < n; ++c)
            for (int i = 0; i < pp; ++i) {
                double s = 0.0;
                for (int kk = 0; kk < k; ++kk)
                    for (int j = 0; j <= i; ++j)
                        s += Aq[((long)r * k + kk) * pp + j] *
                             Bq[((long)kk * n + c) * pp + (i - j)];
                C1q[((long)r * n + c) * pp + i] = s;
            }
    return 0;
}

/* renormhost (big = 0: entry (i, j) at A[(i*n + j)*p + k]) and
   renormhostbig (big = 1: at A[i*n*p + k*n + j]), C++/TCcuda.cu:374-398, in
   place: per entry, k = 0 .. p-2, hi = x_k + x_{k+1}, x_{k+1} = hi - x_k,
   x_k = hi. */
int orc_renormhost(double *A, int n, int p, int big)
{
    if (n < 0 || p < 1) return 1;
    for (long i = 0; i < n; ++i)
        for (long j = 0; j < n; ++j) {
            double *x = big ? A + i * n * p + j : A + (i * n + j) * p;
            const long st = big ? n : 1;
            for (int k = 0; k < p - 1; ++k) {
                const double hi = x[k * st] + x[(k + 1) * st];
                const double lo = hi - x[k * st];
                x[k * st] = hi;
                x[(k + 1) * st] = lo;
            }
        }
    return 0;
}

/* bigB2 layout, C++/TCcuda.cu:79-95: an (n*pp) x (n*pp) column-major matrix,
   element (row = r*pp + j, col = c*pp + i) = Bq[r][c][i-j] for j <= i. */
int orc_bigB2(int n, int pp, const double *Bq, double *out)
{
    long idx = 0;
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < pp; ++i)
            for (int r = 0; r < n; ++r)
                for (int j = 0; j < pp; ++j)
                    out[idx++] = (j <= i) ? Bq[((long)r * n + c) * pp + (i - j)]
                                          : 0.0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* L2: the bit splitter, C++/floatybits.cu                              */
/* ------------------------------------------------------------------ */

/* bitform (floatybits.cu:54-68) followed by double_rep (:92-102) of a copy
   whose 52 stored fraction bits are masked.  The stored fraction is the frexp
   fraction's bits 2^-1 .. 2^-52 (its 2^-53 bit is dropped, :97-99); a
   negative input stores no fraction bits (expandfracbits' loop never runs,
   :20-28); an exponent field <= 0 is stored as 0 (expandexpbits, :40-49). */
typedef struct {
    int sign;          /* stored sign bit */
    int expo;          /* stored exponent field minus 1023 */
    uint64_t frac;     /* 52 stored fraction bits, bit 51 <-> 2^-1 */
} bitrep;

static bitrep bitform(double x)
{
    bitrep b;
    int ex = 0;
    double fr = frexp(x, &ex);
    b.sign = (x >= 0) ? 0 : 1;
    int field = ex + 1023;
    b.expo = (field > 0 ? field : 0) - 1023;
    b.frac = 0;
    if (fr > 0) {
        /* fr in [0.5, 1): the integer fr * 2^53 holds its 53 bits */
        uint64_t m = (uint64_t)ldexp(fr, 53);
        b.frac = m >> 1;
    }
    return b;
}

static double double_rep(bitrep b, uint64_t mask)
{
    double fraction = (double)(b.frac & mask) * 0x1p-52;
    double sgn = b.sign ? -1.0 : 1.0;
    return sgn * fraction * ldexp(1.0, b.expo);
}

/* mask of stored fraction bits [lo, hi) counted from the top (bit 0 <-> 2^-1) */
static uint64_t frac_mask(int lo, int hi)
{
    uint64_t m = 0;
    for (int i = lo; i < hi; ++i)
        m |= (uint64_t)1 << (51 - i);
    return m;
}

/* split4, floatybits.cu:159-191: 13-bit parts [0,13) [13,26) [26,39), the
   remainder is the full stored value minus their sum. */
static void split4(double x, double *d)
{
    bitrep b = bitform(x);
    double d1 = double_rep(b, frac_mask(0, 13));
    double d2 = double_rep(b, frac_mask(13, 26));
    double d3 = double_rep(b, frac_mask(26, 39));
    double full = double_rep(b, frac_mask(0, 52));
    d[0] = d1;
    d[1] = d2;
    d[2] = d3;
    d[3] = full - (d1 + d2 + d3);
}

/* split8, floatybits.cu:218-301: 7,7,7,7,6,6,12-bit parts plus remainder. */
static void split8(double x, double *d)
{
    static const int cut[8] = {0, 7, 14, 21, 28, 34, 40, 52};
    bitrep b = bitform(x);
    double s = 0.0;
    for (int t = 0; t < 7; ++t) {
        d[t] = double_rep(b, frac_mask(cut[t], cut[t + 1]));
    }
    s = d[0];
    for (int t = 1; t < 7; ++t) s = s + d[t];
    d[7] = double_rep(b, frac_mask(0, 52)) - s;
}

/* is_balanced / balance, floatybits.cu:122-148 */
static void balance(double *x, double *y, int threshold)
{
    if (*x == 0.0) return;
    if (*y == 0.0) return;
    int xe, ye;
    frexp(*x, &xe);
    frexp(*y, &ye);
    if (!(xe - ye > threshold - 1)) return;
    double bit = ldexp(1.0, xe - threshold);
    *x = *x - bit;
    *y = *y + bit;
}

/* split4pd, floatybits.cu:196-206 with balance4 :150-154 (pairs a..b-1 with
   b = 4p-2, i.e. the last adjacent pair is never balanced). */
int orc_split4pd(long count, int p, const double *x, double *out)
{
    if (p < 1 || count % p) return 1;
    for (long g = 0; g < count / p; ++g) {
        double *o = out + g * 4 * p;
        for (int l = 0; l < p; ++l)
            split4(x[g * p + l], o + 4 * l);
        for (int i = 0; i < 4 * p - 2; ++i)
            balance(&o[i], &o[i + 1], 14);
    }
    return 0;
}

/* split8pd, floatybits.cu:303-313 with balance8 :208-216 (pairs a..b-2 with
   b = 8p-2; threshold 8 for the first four of every eight, else 7). */
int orc_split8pd(long count, int p, const double *x, double *out)
{
    if (p < 1 || count % p) return 1;
    for (long g = 0; g < count / p; ++g) {
        double *o = out + g * 8 * p;
        for (int l = 0; l < p; ++l)
            split8(x[g * p + l], o + 8 * l);
        for (int i = 0; i < 8 * p - 3; ++i)
            balance(&o[i], &o[i + 1], (i % 8 < 4) ? 8 : 7);
    }
    return 0;
}

/* mixsplit2, floatybits.cu:327-342 with mixbalance :315-325: the high limb
   of a double-double in 4 parts, the low limb in 8; pairs 0..8 balanced with
   thresholds 14 (0-3), 8 (4-7), 7 (8); mixbalance's loop stops at b-2 with
   b = 10, so nine pairs are balanced. */
int orc_mixsplit2(long count, const double *x, double *out)
{
    if (count % 2) return 1;
    for (long g = 0; g < count / 2; ++g) {
        double *o = out + g * 12;
        split4(x[2 * g], o);
        split8(x[2 * g + 1], o + 4);
        for (int i = 0; i < 9; ++i) {
            int t = i % 12;
            balance(&o[i], &o[i + 1], t < 4 ? 14 : (t < 8 ? 8 : 7));
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Seeded, counter-based generators (one independent stream per entry) */
/* ------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t rng_next(uint64_t *s)
{
    *s += 0x9E3779B97F4A7C15ULL;
    return mix64(*s);
}

static inline double rng_u01(uint64_t *s)
{
    return (double)(rng_next(s) >> 11) * 0x1p-53;
}

static inline uint64_t rng_entry(uint64_t seed, uint64_t idx)
{
    return mix64(seed + 0x632BE59BD9B4E019ULL * (idx + 1));
}

/* Fill `count` L-limb entries.
   kind 0: hi ~ U[-1,1), lo = hi*2^-53*U[-0.5,0.5), QuickTwoSum(hi, lo);
           further limbs (L > 2) continue lo's recurrence.
   kind 1: leading limb s*2^e, e uniform in [emin, emax], s ~ +-U[1,2);
           limb i+1 = limb i * 2^-53 * U[-0.5,0.5).
   kind 2: the reference's random_pd (floatybits.cu:107-120): positive,
           exponent field hiexp - 52 i for limb i, 52 random fraction bits
           with the 2^-1 bit set; hiexp = 1023+emin (emin == emax) else
           1023 + emin + (rand mod (emax - emin)). */
int orc_generate(int kind, int L, long count, uint64_t seed, int emin,
                 int emax, double *out)
{
    if (L < 1 || L > 64 || emax < emin) return 1;
    for (long i = 0; i < count; ++i) {
        uint64_t s = rng_entry(seed, (uint64_t)i);
        double *o = out + i * L;
        if (kind == 0) {
            double hi = 2.0 * rng_u01(&s) - 1.0;
            double lo = hi * 0x1p-53 * (rng_u01(&s) - 0.5);
            double e;
            o[0] = qts(hi, lo, &e);
            if (L > 1) o[1] = e;
            for (int l = 2; l < L; ++l)
                o[l] = o[l - 1] * 0x1p-53 * (rng_u01(&s) - 0.5);
        } else if (kind == 1) {
            int span = emax - emin + 1;
            int ex = emin + (int)(rng_next(&s) % (uint64_t)span);
            double sg = (rng_next(&s) & 1) ? -1.0 : 1.0;
            o[0] = sg * ldexp(1.0 + rng_u01(&s), ex);
            for (int l = 1; l < L; ++l)
                o[l] = o[l - 1] * 0x1p-53 * (rng_u01(&s) - 0.5);
        } else if (kind == 2) {
            int hiexp = (emax == emin)
                            ? emin + 1023
                            : emin + 1023 + (int)(rng_next(&s) % (uint64_t)(emax - emin));
            for (int l = 0; l < L; ++l) {
                uint64_t f = (rng_next(&s) >> 12) | ((uint64_t)1 << 51);
                o[l] = (double)f * 0x1p-52 * ldexp(1.0, hiexp - 52 * l - 1023);
            }
        } else {
            return 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Exact checker: two's-complement fixed-point big integers.           */
/* value = sum_w word[w] * 2^(64 w) * 2^BASE, NW words.                 */
/* Every double is m * 2^e with |m| < 2^53, e >= -1074, so products of  */
/* two doubles are integers times 2^(>= -2148) and fit the window       */
/* [2^-2176, 2^2112).                                                   */
/* ------------------------------------------------------------------ */
#define BIG_BASE (-2176)
#define BIG_NW 70

typedef struct {
    uint64_t w[BIG_NW];
} bigfix;

static void big_zero(bigfix *b)
{
    memset(b->w, 0, sizeof(b->w));
}

/* add (neg ? -1 : 1) * mag * 2^shift, shift >= 0 */
static void big_add_u128(bigfix *b, unsigned __int128 mag, int shift, int neg)
{
    if (mag == 0) return;
    int wi = shift >> 6, bi = shift & 63;
    uint64_t part[3];
    part[0] = (uint64_t)(mag << bi);
    part[1] = (uint64_t)((bi ? (mag >> (64 - bi)) : (mag >> 64)));
    part[2] = bi ? (uint64_t)((mag >> 64) >> (64 - bi)) : 0;
    if (!neg) {
        unsigned __int128 carry = 0;
        for (int t = 0; t < 3 && wi + t < BIG_NW; ++t) {
            unsigned __int128 s = (unsigned __int128)b->w[wi + t] + part[t] + carry;
            b->w[wi + t] = (uint64_t)s;
            carry = s >> 64;
        }
        for (int t = wi + 3; carry && t < BIG_NW; ++t) {
            unsigned __int128 s = (unsigned __int128)b->w[t] + carry;
            b->w[t] = (uint64_t)s;
            carry = s >> 64;
        }
    } else {
        uint64_t borrow = 0;
        for (int t = 0; t < 3 && wi + t < BIG_NW; ++t) {
            uint64_t x = b->w[wi + t];
            uint64_t y = part[t];
            uint64_t d = x - y - borrow;
            borrow = (x < y) || (x - y < borrow);
            b->w[wi + t] = d;
        }
        for (int t = wi + 3; borrow && t < BIG_NW; ++t) {
            uint64_t x = b->w[t];
            b->w[t] = x - 1;
            borrow = (x == 0);
        }
    }
}

/* decompose a finite double: x = m * 2^e, m a signed 53-bit integer */
static inline void decomp(double x, int64_t *m, int *e)
{
    if (x == 0.0) {
        *m = 0;
        *e = 0;
        return;
    }
    int ex;
    double fr = frexp(x, &ex);                 /* |fr| in [0.5,1) */
    int64_t mm = (int64_t)ldexp(fr, 53);
    int z = __builtin_ctzll((uint64_t)(mm < 0 ? -mm : mm));
    *m = mm / ((int64_t)1 << z);               /* odd significand */
    *e = ex - 53 + z;                          /* >= -1074 */
}

static void big_add_prod(bigfix *b, int64_t ma, int ea, int64_t mb, int eb)
{
    if (ma == 0 || mb == 0) return;
    int neg = (ma < 0) != (mb < 0);
    unsigned __int128 mag = (unsigned __int128)(uint64_t)(ma < 0 ? -ma : ma) *
                            (uint64_t)(mb < 0 ? -mb : mb);
    big_add_u128(b, mag, ea + eb - BIG_BASE, neg);
}

static void big_add_double(bigfix *b, double x, int negate)
{
    int64_t m;
    int e;
    decomp(x, &m, &e);
    if (m == 0) return;
    int neg = (m < 0) != (negate != 0);
    big_add_u128(b, (unsigned __int128)(uint64_t)(m < 0 ? -m : m), e - BIG_BASE,
                 neg);
}

/* nearest-ish double of the big integer (relative error < 2^-60) */
static double big_to_double(const bigfix *b)
{
    bigfix t = *b;
    int neg = (t.w[BIG_NW - 1] >> 63) != 0;
    if (neg) {
        /* two's complement negate */
        uint64_t carry = 1;
        for (int i = 0; i < BIG_NW; ++i) {
            uint64_t v = ~t.w[i] + carry;
            carry = (carry && v == 0) ? 1 : 0;
            t.w[i] = v;
        }
    }
    int top = BIG_NW - 1;
    while (top >= 0 && t.w[top] == 0) --top;
    if (top < 0) return 0.0;
    double hi = (double)t.w[top];
    double lo = top > 0 ? (double)t.w[top - 1] : 0.0;
    double lo2 = top > 1 ? (double)t.w[top - 2] : 0.0;
    double v = ldexp(hi, 64 * top + BIG_BASE) + ldexp(lo, 64 * (top - 1) + BIG_BASE) +
               ldexp(lo2, 64 * (top - 2) + BIG_BASE);
    return neg ? -v : v;
}

/* Exact value of sum_k A[r][k] * B[k][c] (all limbs), as words. */
static void exact_entry(int L, int k, const double *A, long lda, const double *B,
                        long ldb, int r, int c, bigfix *acc)
{
    big_zero(acc);
    for (int kk = 0; kk < k; ++kk) {
        const double *a = A + ((long)r * lda + kk) * L;
        const double *b = B + ((long)kk * ldb + c) * L;
        int64_t ma[MAXL], mb[MAXL];
        int ea[MAXL], eb[MAXL];
        for (int l = 0; l < L; ++l) {
            decomp(a[l], &ma[l], &ea[l]);
            decomp(b[l], &mb[l], &eb[l]);
        }
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j)
                big_add_prod(acc, ma[i], ea[i], mb[j], eb[j]);
    }
}

/* For each sampled entry (rows[s], cols[s]):
     err[s] = exact(C) - sum of the L limbs of C_hat   (rounded to double)
     den[s] = sum_k |A[r][k]| * |B[k][c]|              (leading limbs, double)
     val[s] = exact(C)                                  (rounded to double)
   Rejects non-finite inputs (returns 2). */
int orc_exact_error(int L, int m, int n, int k, const double *A, long lda,
                    const double *B, long ldb, const double *C, long ldc,
                    int nsamp, const int *rows, const int *cols, double *err,
                    double *den, double *val, int nthreads)
{
    if (L < 1 || L > MAXL) return 1;
    int bad = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(| : bad)
    for (int s = 0; s < nsamp; ++s) {
        int r = rows[s], c = cols[s];
        if (r < 0 || r >= m || c < 0 || c >= n) {
            bad |= 1;
            continue;
        }
        bigfix acc;
        exact_entry(L, k, A, lda, B, ldb, r, c, &acc);
        val[s] = big_to_double(&acc);
        const double *ch = C + ((long)r * ldc + c) * L;
        for (int l = 0; l < L; ++l) {
            if (!isfinite(ch[l])) bad |= 2;
            else big_add_double(&acc, ch[l], 1);
        }
        err[s] = big_to_double(&acc);
        double d = 0.0;
        for (int kk = 0; kk < k; ++kk)
            d += fabs(A[((long)r * lda + kk) * L]) * fabs(B[((long)kk * ldb + c) * L]);
        den[s] = d;
    }
    return bad ? (bad & 1 ? 1 : 2) : 0;
}

/* The exact value of one entry as BIG_NW little-endian two's-complement
   words scaled by 2^BIG_BASE; used to pin the checker against exact rational
   arithmetic. */
int orc_exact_entry_words(int L, int k, const double *A, long lda,
                          const double *B, long ldb, int r, int c,
                          uint64_t *words, int *base, int *nwords)
{
    bigfix acc;
    exact_entry(L, k, A, lda, B, ldb, r, c, &acc);
    memcpy(words, acc.w, sizeof(acc.w));
    *base = BIG_BASE;
    *nwords = BIG_NW;
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
