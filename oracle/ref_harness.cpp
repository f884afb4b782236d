// ref_harness.cpp -- binds the reference repository's own host-compilable
// arithmetic and splitter (C++/{double,quad,octo,hexa}_double_functions.cu,
// C++/floatybits.cu, compiled from where they lie by oracle/Makefile) to a
// flat extern "C" interface, so the C restatement in md_oracle.c can be
// checked against them bit for bit.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/libmdref.so, never linked
// into the product.  Nothing here is copied from the reference: the GEMM and
// squeeze loops below are this project's own drivers around the reference's
// scalar functions, following the loop order of C++/TCcuda.cu:320-372 and
// C++/TCcuda.cu:876-906.
#include <cstring>
#include <vector>

#include "double_double_functions.h"
#include "quad_double_functions.h"
#include "octo_double_functions.h"
#include "hexa_double_functions.h"
#include "floatybits.h"

#define V4(a) a[0], a[1], a[2], a[3]
#define V8(a) V4(a), a[4], a[5], a[6], a[7]
#define V16(a) V8(a), a[8], a[9], a[10], a[11], a[12], a[13], a[14], a[15]
#define V5(a) V4(a), a[4]
#define V9(a) V8(a), a[8]
#define V17(a) V16(a), a[16]
#define P4(a) &a[0], &a[1], &a[2], &a[3]
#define P8(a) P4(a), &a[4], &a[5], &a[6], &a[7]
#define P16(a) P8(a), &a[8], &a[9], &a[10], &a[11], &a[12], &a[13], &a[14], &a[15]

static void ref_mul1(int L, const double *a, const double *b, double *c)
{
    switch (L) {
    case 2: ddf_mul(a[0], a[1], b[0], b[1], &c[0], &c[1]); break;
    case 4: qdf_mul(V4(a), V4(b), P4(c)); break;
    case 8: odf_mul(V8(a), V8(b), P8(c)); break;
    case 16: hdf_mul(V16(a), V16(b), P16(c)); break;
    }
}

static void ref_add1(int L, const double *a, const double *b, double *c)
{
    switch (L) {
    case 2: ddf_add(a[0], a[1], b[0], b[1], &c[0], &c[1]); break;
    case 4: qdf_add(V4(a), V4(b), P4(c)); break;
    case 8: odf_add(V8(a), V8(b), P8(c)); break;
    case 16: hdf_add(V16(a), V16(b), P16(c)); break;
    }
}

static void ref_inc_d1(int L, double *a, double y)
{
    switch (L) {
    case 2: ddf_inc_d(&a[0], &a[1], y); break;
    case 4: qdf_inc_d(P4(a), y); break;
    case 8: odf_inc_d(P8(a), y); break;
    case 16: hdf_inc_d(P16(a), y); break;
    }
}

static bool ok_L(int L) { return L == 2 || L == 4 || L == 8 || L == 16; }

// The reference's off-path scalar operations, one entry; op codes as in
// md_oracle.c (0 add, 1 sub, 2 mul, 3 div, 4 sqr, 5 sqrt, 6 mul by a double,
// 7 inc_d, 8 mul_pwr2, 9 abs, 10 minus).
static int ref_op1(int op, int L, const double *a, const double *b, double *c)
{
    double t[16];
    std::memcpy(t, a, sizeof(double) * L);
    switch (op) {
    case 0: ref_add1(L, a, b, c); return 0;
    case 2: ref_mul1(L, a, b, c); return 0;
    case 7:
        ref_inc_d1(L, t, b[0]);
        std::memcpy(c, t, sizeof(double) * L);
        return 0;
    case 10:
        switch (L) {
        case 2: ddf_minus(&t[0], &t[1]); break;
        case 4: qdf_minus(P4(t)); break;
        case 8: odf_minus(P8(t)); break;
        case 16: hdf_minus(P16(t)); break;
        }
        std::memcpy(c, t, sizeof(double) * L);
        return 0;
    }
    switch (L) {
    case 2:
        switch (op) {
        case 1: ddf_sub(a[0], a[1], b[0], b[1], &c[0], &c[1]); return 0;
        case 3: ddf_div(a[0], a[1], b[0], b[1], &c[0], &c[1]); return 0;
        case 4: ddf_sqr(a[0], a[1], &c[0], &c[1]); return 0;
        case 5: ddf_sqrt(a[0], a[1], &c[0], &c[1]); return 0;
        case 6: ddf_mul_d_dd(b[0], a[0], a[1], &c[0], &c[1]); return 0;
        case 8: c[0] = a[0] * b[0]; c[1] = a[1] * b[0]; return 0;
        case 9: ddf_abs(a[0], a[1], &c[0], &c[1]); return 0;
        }
        break;
    case 4:
        switch (op) {
        case 1: qdf_sub(V4(a), V4(b), P4(c)); return 0;
        case 3: qdf_div(V4(a), V4(b), P4(c)); return 0;
        case 4: qdf_sqr(V4(a), P4(c)); return 0;
        case 5: qdf_sqrt(V4(a), P4(c)); return 0;
        case 6: qdf_mul_qd_d(V4(a), b[0], P4(c)); return 0;
        case 8: qdf_mul_pwr2(V4(a), b[0], P4(c)); return 0;
        case 9: qdf_abs(V4(a), P4(c)); return 0;
        }
        break;
    case 8:
        switch (op) {
        case 1: odf_sub(V8(a), V8(b), P8(c)); return 0;
        case 3: odf_div(V8(a), V8(b), P8(c)); return 0;
        case 4: odf_sqr(V8(a), P8(c)); return 0;
        case 5: odf_sqrt(V8(a), P8(c)); return 0;
        case 6: odf_mul_od_d(V8(a), b[0], P8(c)); return 0;
        case 8: odf_mul_pwr2(V8(a), b[0], P8(c)); return 0;
        case 9: odf_abs(V8(a), P8(c)); return 0;
        }
        break;
    case 16:
        switch (op) {
        case 1: hdf_sub(V16(a), V16(b), P16(c)); return 0;
        case 3: hdf_div(V16(a), V16(b), P16(c)); return 0;
        case 4: hdf_sqr(V16(a), P16(c)); return 0;
        case 5: hdf_sqrt(V16(a), P16(c)); return 0;
        case 6: hdf_mul_hd_d(V16(a), b[0], P16(c)); return 0;
        case 8: hdf_mul_pwr2(V16(a), b[0], P16(c)); return 0;
        case 9: hdf_abs(V16(a), P16(c)); return 0;
        }
        break;
    }
    return 1;
}

extern "C" {

int ref_mul(int L, long count, const double *a, const double *b, double *c)
{
    if (!ok_L(L)) return 1;
    for (long i = 0; i < count; ++i) ref_mul1(L, a + i * L, b + i * L, c + i * L);
    return 0;
}

int ref_add(int L, long count, const double *a, const double *b, double *c)
{
    if (!ok_L(L)) return 1;
    for (long i = 0; i < count; ++i) ref_add1(L, a + i * L, b + i * L, c + i * L);
    return 0;
}

int ref_inc_d(int L, long count, const double *a, const double *y, double *c)
{
    if (!ok_L(L)) return 1;
    for (long i = 0; i < count; ++i) {
        double t[16];
        std::memcpy(t, a + i * L, sizeof(double) * L);
        ref_inc_d1(L, t, y[i]);
        std::memcpy(c + i * L, t, sizeof(double) * L);
    }
    return 0;
}

int ref_op(int op, int L, long count, const double *a, const double *b, double *c)
{
    if (!ok_L(L) || op < 0 || op > 10) return 1;
    const int bstride = (op == 6 || op == 7 || op == 8) ? 1 : (op <= 3 ? L : 0);
    for (long i = 0; i < count; ++i)
        if (ref_op1(op, L, a + i * L, b ? b + i * bstride : nullptr, c + i * L)) return 1;
    return 0;
}

int ref_fast_renorm(int L, long count, const double *x, double *r)
{
    for (long i = 0; i < count; ++i) {
        const double *f = x + i * (L + 1);
        double *o = r + i * L;
        switch (L) {
        case 4: qdf_fast_renorm(V5(f), P4(o)); break;
        case 8: odf_fast_renorm(V9(f), P8(o)); break;
        case 16: hdf_fast_renorm(V17(f), P16(o)); break;
        default: return 1;
        }
    }
    return 0;
}

int ref_renorm(int L, long count, const double *x, double *r)
{
    for (long i = 0; i < count; ++i) {
        const double *f = x + i * (L + 1);
        double *o = r + i * L;
        double pr;
        switch (L) {
        case 4: qdf_renorm4(V5(f), &pr, P4(o)); break;
        case 8: odf_renorm8(V9(f), &pr, P8(o)); break;
        case 16: hdf_renorm16(V17(f), &pr, P16(o)); break;
        default: return 1;
        }
    }
    return 0;
}

int ref_two_prod(long count, const double *a, const double *b, double *p, double *e)
{
    for (long i = 0; i < count; ++i) p[i] = ddf_two_prod(a[i], b[i], &e[i]);
    return 0;
}

// C[r][c] = fold_k add(acc, mul(A[r][k], B[k][c])) with acc held in C, as the
// reference's host loops do (C++/TCcuda.cu:320-372), for rectangular shapes.
int ref_gemm(int L, int m, int n, int k, const double *A, long lda,
             const double *B, long ldb, double *C, long ldc)
{
    if (!ok_L(L)) return 1;
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < n; ++c) {
            double *acc = C + ((long)r * ldc + c) * L;
            for (int l = 0; l < L; ++l) acc[l] = 0.0;
            for (int kk = 0; kk < k; ++kk) {
                double prod[16];
                ref_mul1(L, A + ((long)r * lda + kk) * L, B + ((long)kk * ldb + c) * L, prod);
                double t[16];
                ref_add1(L, acc, prod, t);
                std::memcpy(acc, t, sizeof(double) * L);
            }
        }
    return 0;
}

// y = fold_j inc_d(y, x[j]), C++/TCcuda.cu:876-906 per entry.
int ref_squeeze(int L, long nent, int pp, const double *x, double *y)
{
    if (!ok_L(L)) return 1;
    for (long i = 0; i < nent; ++i) {
        double acc[16] = {0};
        for (int j = 0; j < pp; ++j) ref_inc_d1(L, acc, x[i * pp + j]);
        std::memcpy(y + i * L, acc, sizeof(double) * L);
    }
    return 0;
}

int ref_split4pd(long count, int p, const double *x, double *out)
{
    std::vector<double> v(x, x + count);
    std::vector<double> r = split4pd(v, p);
    std::memcpy(out, r.data(), sizeof(double) * r.size());
    return (long)r.size() == 4 * count ? 0 : 1;
}

int ref_split8pd(long count, int p, const double *x, double *out)
{
    std::vector<double> v(x, x + count);
    std::vector<double> r = split8pd(v, p);
    std::memcpy(out, r.data(), sizeof(double) * r.size());
    return (long)r.size() == 8 * count ? 0 : 1;
}

int ref_mixsplit2(long count, const double *x, double *out)
{
    std::vector<double> v(x, x + count);
    std::vector<double> r = mixsplit2(v);
    std::memcpy(out, r.data(), sizeof(double) * r.size());
    return (long)r.size() == 6 * count ? 0 : 1;
}

}  // extern "C"
