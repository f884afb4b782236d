/*
 * mdgemm.h -- C ABI of the MI355X (gfx950) multiple-double GEMM library
 * (libmdgemm.so).
 *
 * Numbers are p-doubles: L in {2, 4, 8, 16} FP64 limbs (double, quad, octo
 * and hexa-double), most significant limb first.  Matrices use the
 * reference's interleaved "component-array" layout: entry (r, c) of a
 * row-major matrix with leading dimension ld (in entries) occupies
 * X[(r * ld + c) * L + limb]  (C++/TCcuda.cu:130-131,154-155).
 *
 * Every matrix pointer is a DEVICE pointer (hipMalloc'd by the caller); the
 * library allocates only its own scratch.  `stream` is a hipStream_t (NULL =
 * the null stream); every call is asynchronous on it unless noted.
 * Every entry point returns an mdg_status; mdg_last_error() gives the text
 * of the last failure on the calling thread.  Host threads may call the
 * library at the same time on different streams (scratch is kept per device
 * and stream); calls from several threads on one stream are not supported.
 *
 * The reference has no error reporting at all (an unsupported p silently
 * returns zeros, C++/TCcuda.cu:269-297); here unsupported limb counts, bad
 * shapes, null pointers and non-finite inputs are rejected.
 */
#ifndef MDGEMM_H
#define MDGEMM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mdg_status {
    MDG_SUCCESS = 0,
    MDG_ERROR_INVALID_VALUE = 1,   /* bad L, shape, leading dimension, pointer */
    MDG_ERROR_NOT_SUPPORTED = 2,   /* valid request the library does not serve */
    MDG_ERROR_HIP = 3,             /* a HIP runtime call failed */
    MDG_ERROR_OUT_OF_MEMORY = 4,   /* scratch allocation failed */
    MDG_ERROR_NON_FINITE = 5,      /* an input limb is Inf or NaN */
    MDG_ERROR_RANGE = 6            /* exponent range beyond the slice scheme */
} mdg_status;

/* GEMM methods for mdg_gemm / mdg_gemm_batched */
typedef enum mdg_method {
    MDG_METHOD_VECTOR = 0,      /* one thread per output entry, p-double mul+add */
    MDG_METHOD_F64_SPLIT = 1,   /* signed fixed-point slices on FP64 MFMA */
    MDG_METHOD_INT8_OZAKI = 2,  /* 8-bit slices on INT8 MFMA, INT32 accumulate */
    MDG_METHOD_BF16_OZAKI = 3,  /* 8-bit slices on BF16 MFMA, FP32 accumulate */
    MDG_METHOD_INT8_RESIDUE = 4, /* INT8 residues modulo coprime moduli on INT8
                                    MFMA, exact CRT reconstruction (no
                                    counterpart in the reference) */
    MDG_METHOD_F64_RESIDUE = 5   /* residues modulo primes below 2^21 on FP64
                                    MFMA, exact CRT reconstruction (no
                                    counterpart in the reference) */
} mdg_method;

/* two_prod flavour of the vector path */
typedef enum mdg_twoprod {
    MDG_TWOPROD_FMA = 0,     /* p = a*b, e = fma(a,b,-p) */
    MDG_TWOPROD_DEKKER = 1   /* the reference's Veltkamp/Dekker product */
} mdg_twoprod;

/* What a split-product GEMM did (filled when a non-NULL pointer is given). */
typedef struct mdg_gemm_info {
    int method;            /* mdg_method actually run */
    int slices;            /* S: slices per operand entry (residue method:
                              moduli) */
    int slice_bits;        /* w: bits per slice (residue method: 8) */
    int diagonals;         /* slice-sum diagonals kept (== slices; residue
                              method: wide moduli, three GEMMs each) */
    long long pair_gemms;  /* P: slice-pair products issued, S(S+1)/2 (residue
                              method: INT8 GEMMs of depth k, wide moduli
                              counted 4 times) */
    int total_bits;        /* W fixed-point bits per entry (S * w for slices) */
    int spread_a;          /* exponent spread estimate of A, bits */
    int spread_b;          /* exponent spread estimate of B, bits */
    int flushed;           /* 1 if accumulators were widened inside the K loop */
    /* stage times in ms, measured with HIP events on the call's stream when
       profiling is on (mdg_set_profiling), else 0:
       [0] exponent scan, [1] slicing, [2] diagonal GEMMs, [3] reconstruction */
    float stage_ms[4];
    int backend;           /* GEMM engine that ran: 2 this library's MFMA
                              kernels (band kernels for slices), 4 a
                              single-diagonal slice kernel (256 x 256 INT8,
                              INT64-plane BF16), 1 rocBLAS, 3 hipBLASLt
                              (INT8 only) */
} mdg_gemm_info;

/* One problem of a batch (all pointers on the device). */
typedef struct mdg_problem {
    int L, m, n, k;
    const double *A;
    long long lda;
    const double *B;
    long long ldb;
    double *C;
    long long ldc;
} mdg_problem;

const char *mdg_last_error(void);
int mdg_version(void);

/* Initialise the library on the current device (optional; every entry point
   initialises on first use). */
int mdg_init(void);

/* Release cached scratch buffers and cached batch plans (their graphs and
   workspaces) of the current device; waits for the streams that use them. */
int mdg_release_workspace(void);

/* Per-stage HIP-event timing of mdg_gemm (on = 1) into mdg_gemm_info.stage_ms;
   when on, mdg_gemm waits for its own last event before returning. */
int mdg_set_profiling(int on);

/* Engine of the INT8 GEMMs (slice planes and residue products) and of the
   BF16 slice planes: 0 or 2 = this library's MFMA kernels (the default: the
   band kernels for INT8 and BF16 slice planes, the plain INT8 GEMM for
   residue products), 1 = rocBLAS, 3 = hipBLASLt (INT8 only; the first call
   of a shape times the algorithms it proposes and keeps the fastest),
   4 = the single-diagonal kernels for slice planes (256 x 256 INT8,
   INT64-plane BF16; residue products as 0).  Engines 1, 3 and 4 are there
   for comparison.  All are exact and give identical results; they differ
   only in speed.  Process-wide. */
int mdg_set_int8_backend(int backend);

/* ----------------------------------------------------------------------
 * Vector GEMM.  C = A * B with one thread per output entry, each a k-step
 * chain acc = add(acc, mul(A[r][k], B[k][c])), acc starting at zero, k
 * ascending -- the reference kernels ddmm/qdmm/odmm/hdmm
 * (C++/TCcuda.cu:119-246) and its host loops matmul{dd,qd,od,hd}f
 * (C++/TCcuda.cu:320-372), generalised to m x k times k x n with leading
 * dimensions.  Bit-identical to the reference's host arithmetic with
 * MDG_TWOPROD_DEKKER, and with MDG_TWOPROD_FMA whenever no partial product
 * underflows.
 * -------------------------------------------------------------------- */
int mdg_vector_gemm(int L, int m, int n, int k, const double *A, long long lda,
                    const double *B, long long ldb, double *C, long long ldc,
                    int twoprod, void *stream);

/* ----------------------------------------------------------------------
 * Split-product GEMM C = A * B (no reference counterpart as a single call;
 * it replaces the reference's split -> tensor-core GEMM -> squeeze chain,
 * C++/TCtest.cu:53-114, for signed inputs of any exponent range).
 *   method       MDG_METHOD_F64_SPLIT / _INT8_OZAKI / _BF16_OZAKI /
 *                _INT8_RESIDUE / _F64_RESIDUE, or MDG_METHOD_VECTOR (then
 *                identical to mdg_vector_gemm, FMA).  The _RESIDUE methods
 *                compute the fixed-point product from residues modulo
 *                pairwise coprime moduli (Ozaki scheme II): GEMMs linear,
 *                not quadratic, in the bit count; k <= 2^22.
 *   target_bits  fixed-point precision goal; <= 0 means 53*L.  The slice
 *                count is chosen from it and from the exponent spread of
 *                each operand's rows/columns (a small device->host read).
 * Synchronises `stream` once (to read the exponent spread).
 * -------------------------------------------------------------------- */
int mdg_gemm(int L, int m, int n, int k, const double *A, long long lda,
             const double *B, long long ldb, double *C, long long ldc,
             int method, int target_bits, mdg_gemm_info *info, void *stream);

/* ----------------------------------------------------------------------
 * mdg_gemm without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results (bit for bit) and errors as mdg_gemm, for the methods
 * VECTOR, F64_SPLIT, INT8_OZAKI (always the hand-written INT8 kernels) and
 * BF16_OZAKI; the residue methods return MDG_ERROR_NOT_SUPPORTED.  All
 * workspace lives in one static device array of
 * mdg_noalloc_workspace_bytes() bytes (256 MiB) that every GPU the library
 * runs on holds until the process ends; a problem that needs more returns
 * MDG_ERROR_OUT_OF_MEMORY with the size in mdg_last_error() (C untouched).
 * Every call on one GPU shares that GPU's workspace: calls that overlap in
 * time on one GPU (different streams or host threads) race and must be
 * serialised by the caller.  Like mdg_gemm, it synchronises `stream` once
 * to read the exponent spread.  (Quad-double fits up to about n = 1024.)
 * -------------------------------------------------------------------- */
int mdg_gemm_noalloc(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                     double *C, long long ldc, int method, int target_bits, mdg_gemm_info *info, void *stream);

/* Size in bytes of mdg_gemm_noalloc's static workspace. */
size_t mdg_noalloc_workspace_bytes(void);

/* ----------------------------------------------------------------------
 * Batched GEMM: `count` independent problems (L in {2,4,8,16} may differ
 * per problem), all computed with `method`.  `problems` is a HOST array.
 *   spread_bits  exponent-spread allowance per operand (bits) used to size
 *                the slices without reading the data back; < 0 means 16.
 *   persistent   1: one persistent launch per stage over a device work
 *                list, captured once into a hipGraph and replayed on later
 *                calls with the same batch shape; 0: a loop of per-problem
 *                launches (the baseline).
 * -------------------------------------------------------------------- */
int mdg_gemm_batched(int count, const mdg_problem *problems, int method,
                     int target_bits, int spread_bits, int persistent,
                     void *stream);

/* Slice and residue batches size their slices from spread_bits without
   reading the data back; the scans then check every problem on the device:
   a problem with an Inf/NaN input limb (flag bit 0) or an exponent spread
   of A's rows or B's columns above spread_bits (bit 1, too few slices for
   the precision target) gets C filled with NaN instead of a silently wrong
   result.  mdg_batched_last_flags copies the flags of the last batched call
   on `stream` (count must match it) into flags[count], synchronising the
   stream, and returns MDG_ERROR_NON_FINITE or MDG_ERROR_RANGE for the first
   flagged problem (index in mdg_last_error()), else MDG_SUCCESS. */
int mdg_batched_last_flags(void *stream, int count, int *flags);

/* Duration in ms of the GEMM stage (the one launch of INT8 tiles over every
   problem) of the last replay on `stream` of a fused persistent batch whose
   plan was built while mdg_set_profiling(1) was on (HIP events recorded
   into its graph); synchronises on the stage's end. */
int mdg_batched_last_gemm_ms(void *stream, float *ms);

/* Bounds of the persistent-batch plan cache (process-wide): at most
   max_plans plans and max_bytes of device memory in them; the least
   recently used plans are freed first (defaults 32 plans, 16 GiB).  A plan
   is keyed by the batch's shapes, pointers, method and stream, so a caller
   that passes new buffers on every call builds a new plan each time. */
int mdg_set_batch_plan_cache(int max_plans, long long max_bytes);

/* ----------------------------------------------------------------------
 * The reference's split ("tensor core") path, piece by piece.
 * -------------------------------------------------------------------- */

/* split4pd (C++/floatybits.cu:196-206, balance4 :150-154): each limb of
   `count` doubles (count a multiple of p) becomes 4 parts of 13/13/13/13
   stored fraction bits; adjacent parts of one p-double are exponent-balanced.
   out holds 4*count doubles.  The reference's bit semantics are kept,
   including that a negative limb splits into zero parts (floatybits.cu:20-28)
   and that the 53rd significand bit is dropped (:97-99). */
int mdg_split4pd(const double *x, long long count, int p, double *out, void *stream);

/* split8pd (C++/floatybits.cu:303-313, balance8 :208-216): 8 parts of
   7,7,7,7,6,6,12 bits plus remainder; out holds 8*count doubles. */
int mdg_split8pd(const double *x, long long count, int p, double *out, void *stream);

/* mixsplit2 (C++/floatybits.cu:327-342, mixbalance :315-325): double-doubles,
   high limb in 4 parts and low limb in 8; out holds 6*count doubles. */
int mdg_mixsplit2(const double *x, long long count, double *out, void *stream);

/* bigB2 (C++/TCcuda.cu:79-95): the (n*pp) x (n*pp) column-major block
   Toeplitz expansion of an n x n matrix of pp-part entries.  Provided for
   callers that want the explicit operand; mdg_conv_gemm never builds it. */
int mdg_bigB2(const double *Bq, int n, int pp, double *out, void *stream);

/* matmulhost (C++/TCcuda.cu:304-317): C[r][c][i] = sum_k sum_{j<=i}
   A[r][k][j] * B[k][c][i-j] for n x n matrices of p-part entries, in plain
   doubles with every product and sum rounded, k ascending and j = 0 .. i
   within it (the reference's host loop order, so the same bits). */
int mdg_matmulhost(int p, int n, const double *A, const double *B, double *C, void *stream);

/* renormhost (big = 0, interleaved entries A[(i*n + j)*p + k]) and
   renormhostbig (big = 1, A[i*n*p + k*n + j], the bigA layout),
   C++/TCcuda.cu:374-398, in place: per entry, k = 0 .. p-2 in turn,
   hi = x_k + x_{k+1}; x_{k+1} = hi - x_k; x_k = hi. */
int mdg_renormhost(double *A, int n, int p, int big, void *stream);

/* The reference's FP64 tensor-core kernel `matmul` (C++/TCtest.cu:21-46):
   C (M x N, row-major) += A (M x K, row-major) * B (K x N, column-major), on
   FP64 MFMA.  Accumulates into C like the reference. */
int mdg_dgemm_acc(int M, int N, int K, const double *A, const double *B, double *C,
                  void *stream);

/* Limb convolution C1q = Aq * bigB2(Bq) without materialising bigB2:
   C1q[r][c][i] = sum_k sum_{j<=i} Aq[r][k][j] * Bq[k][c][i-j] for an
   m x k and a k x n matrix of pp-part entries (C++/TCcuda.cu:79-95 with
   C++/TCtest.cu:21-46, square n in the reference).  FP64 MFMA. */
int mdg_conv_gemm(int m, int n, int k, int pp, const double *Aq, const double *Bq,
                  double *C1q, void *stream);

/* pllsqueeze_old (C++/TCcuda.cu:876-958): y[i] = fold_j inc_d(y[i], x[i*pp+j])
   from zero, for `nent` entries of pp parts into L-limb numbers. */
int mdg_squeeze(int L, long long nent, int pp, const double *x, double *y, void *stream);

/* The whole reference split path for L-limb n x n inputs, fused on the
   device: split (split_kind 4 -> split4pd, 8 -> split8pd, 12 -> mixsplit2,
   L == 2 only), limb convolution on FP64 MFMA, squeeze into L limbs.
   Equal to the reference's C++/TCtest.cu:53-114 chain wherever the
   reference's FP64 accumulation is exact (inputs of equal leading
   exponent, positive entries: its own test domain). */
int mdg_tc_gemm(int L, int m, int n, int k, const double *A, const double *B,
                double *C, int split_kind, void *stream);

/* ----------------------------------------------------------------------
 * Elementwise p-double arithmetic (the reference's scalar library, off the
 * GEMM path): C++/double_double_functions.cu, C++/quad_double_functions.cu,
 * C++/octo_double_functions.cu, C++/hexa_double_functions.cu.
 * -------------------------------------------------------------------- */

typedef enum mdg_op {
    MDG_OP_ADD = 0,       /* c = a + b   Xdf_add (quad :216-242)           */
    MDG_OP_SUB = 1,       /* c = a - b   Xdf_sub (double :61-73, quad :299-309) */
    MDG_OP_MUL = 2,       /* c = a * b   Xdf_mul (quad :321-381)           */
    MDG_OP_DIV = 3,       /* c = a / b   Xdf_div (double :230-248, quad :485-520) */
    MDG_OP_SQR = 4,       /* c = a * a   Xdf_sqr (double :208-216, quad :383-440) */
    MDG_OP_SQRT = 5,      /* c = sqrt(a) Xdf_sqrt (double :250-279, quad :524-537);
                             a negative leading limb gives the reference's -1
                             root at double-double and whatever its Heron
                             step then makes of it above */
    MDG_OP_MUL_D = 6,     /* c = a * d   ddf_mul_d_dd (double :218-228),
                             Xdf_mul_X_d (quad :442-468)                  */
    MDG_OP_INC_D = 7,     /* c = a + d   Xdf_inc_d (double :99-106, quad :269-275) */
    MDG_OP_MUL_PWR2 = 8,  /* c = a * d, d a power of two: limb-wise
                             (quad :311-319)                             */
    MDG_OP_ABS = 9,       /* c = |a|     Xdf_abs (quad :194-212); at L = 16
                             the reference copies limb 5 into limb 9
                             (hexa :1479,1489), this returns |a| limb-wise */
    MDG_OP_NEG = 10       /* c = -a      Xdf_minus (quad :290-297)         */
} mdg_op;

/* c[i] = op(a[i], b[i]) for n entries of L-limb numbers in the interleaved
   layout (entry i at [i * L + limb]).  b: n L-limb entries for ADD, SUB, MUL
   and DIV; n doubles (b[i]) for MUL_D, INC_D and MUL_PWR2; ignored (may be
   NULL) for the unary ops.  c may alias a or b entry for entry.  twoprod
   selects the exact product (mdg_twoprod): MDG_TWOPROD_DEKKER gives the
   reference's host bits; the two agree wherever no partial product
   underflows.  Division by zero and roots of zero (L >= 4) give Inf/NaN as
   the reference's do; no input is rejected for its value. */
int mdg_elementwise(int op, int L, long long n, const double *a, const double *b, double *c, int twoprod,
                    void *stream);

/* ----------------------------------------------------------------------
 * Utilities.
 * -------------------------------------------------------------------- */

/* Limb-planar copy of an interleaved rows x cols matrix (row stride ldx
   entries): planar[l * rows * cols + r * cols + c] = X[(r * ldx + c) * L + l],
   every limb's part() (C++/TCcuda.cu:34-40) at once.  The buffers must not
   overlap. */
int mdg_to_planar(int L, int rows, int cols, const double *x, long long ldx, double *planar, void *stream);

/* The inverse of mdg_to_planar. */
int mdg_from_planar(int L, int rows, int cols, const double *planar, double *x, long long ldx, void *stream);

/* Y = X^T for an interleaved rows x cols matrix X (row stride ldx); Y is
   cols x rows with row stride ldy >= rows.  The GEMMs take row strides but
   no transposes (the reference is square and row-major only): a transposed
   operand is materialised with this first.  The buffers must not overlap. */
int mdg_transpose(int L, int rows, int cols, const double *x, long long ldx, double *y, long long ldy,
                  void *stream);

/* Seeded counter-based generator: `count` L-limb entries.
   kind 0: hi ~ U[-1,1), lo = hi*2^-53*U[-0.5,0.5), QuickTwoSum(hi, lo)
   kind 1: leading limb s*2^e, e uniform in [emin, emax], s ~ +-U[1,2);
           next limb = previous * 2^-53 * U[-0.5,0.5)
   kind 2: the reference's random_pd layout (C++/floatybits.cu:107-120):
           positive limbs of 52 fraction bits at exponents hiexp - 52 i. */
int mdg_fill_random(int kind, int L, long long count, unsigned long long seed,
                    int emin, int emax, double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MDGEMM_H */
