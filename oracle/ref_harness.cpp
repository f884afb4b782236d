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
