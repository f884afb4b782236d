// api.hip -- the extern "C" entry points of libmdgemm.so (include/mdgemm.h):
// argument validation, status codes, scratch management and dispatch.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"

namespace mdg {

int vector_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                double *C, long long ldc, int twoprod, hipStream_t s);
int split_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb, double *C,
               long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info, hipStream_t s);
int dgemm_acc(int M, int N, int K, const double *A, const double *B, double *C, hipStream_t s);
int conv_gemm(int m, int n, int k, int pp, const double *Aq, const double *Bq, double *C1q, hipStream_t s);
int conv_squeeze(int L, int m, int n, int k, int pp, const double *Aq, const double *Bq, double *C, hipStream_t s);
int ref_split(int kind, const double *x, long long count, int p, double *out, hipStream_t s);
int squeeze(int L, long long nent, int pp, const double *x, double *y, hipStream_t s);
int elementwise(int op, int L, long long n, const double *a, const double *b, double *c, int tp, hipStream_t s);
bool elementwise_binary(int op);
bool elementwise_takes_double(int op);
int to_planar(int L, int rows, int cols, const double *x, long long ldx, double *planar, hipStream_t s);
int from_planar(int L, int rows, int cols, const double *planar, double *x, long long ldx, hipStream_t s);
int transpose(int L, int rows, int cols, const double *x, long long ldx, double *y, long long ldy, hipStream_t s);
int bigB2(const double *Bq, int n, int pp, double *out, hipStream_t s);
int matmulhost(const double *A, const double *B, int n, int p, double *C, hipStream_t s);
int renormhost(double *A, int n, int p, int big, hipStream_t s);
int fill_random(int kind, int L, long long count, unsigned long long seed, int emin, int emax, double *out,
                hipStream_t s);
int batched_gemm(int count, const mdg_problem *problems, int method, int target_bits, int spread_bits,
                 int persistent, hipStream_t s);
int noalloc_workspace(void **ptr);
int release_batch_plans();
int set_batch_plan_cache(int max_plans, long long max_bytes);
int batched_last_flags(void *stream, int count, int *out);
int batched_last_gemm_ms(void *stream, float *ms);

static thread_local char g_err[512] = "";
static std::atomic<int> g_profiling{0};

bool arena_active();

bool profiling_on() { return g_profiling.load(std::memory_order_relaxed) != 0 && !arena_active(); }

static std::atomic<int> g_int8_backend{0};

int int8_backend() { return g_int8_backend.load(std::memory_order_relaxed); }

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *get_error() { return g_err; }

// ------------------------------------------------------------ scratch --
struct ScratchBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

static std::mutex g_scratch_mu;
static std::map<std::pair<int, void *>, std::vector<ScratchBuf>> g_scratch;

// The arena installed on the calling thread (mdg_gemm_noalloc), if any
static thread_local Arena *g_arena = nullptr;

void arena_begin(void *base, size_t cap, Arena *a)
{
    *a = Arena{};
    a->base = (char *)base;
    a->cap = cap;
    g_arena = a;
}

void arena_end() { g_arena = nullptr; }

void arena_reset()
{
    if (g_arena) g_arena->used = 0;
}

bool arena_active() { return g_arena != nullptr; }

bool noalloc_active() { return g_arena != nullptr && g_arena->noalloc; }

// while set, scratch() (allocating path) adds every request's size, rounded
// to 256 bytes plus 256 of alignment slack, to *g_rec
static thread_local size_t *g_rec = nullptr;

void scratch_record(size_t *counter) { g_rec = counter; }

bool arena_overflow(size_t *need)
{
    if (!g_arena || !g_arena->overflow) return false;
    *need = g_arena->need;
    return true;
}

void *scratch(void *stream, size_t bytes, int slot)
{
    if (g_arena) {
        Arena &a = *g_arena;
        // pieces start at 256-byte aligned device addresses
        const uintptr_t at = (uintptr_t)a.base + a.used;
        const size_t off = (size_t)(((at + 255) & ~(uintptr_t)255) - (uintptr_t)a.base);
        if (a.overflow || off + bytes > a.cap) {
            // keep counting what the whole call would need
            a.need = std::max(a.need, off) + bytes;
            a.used = a.need;
            a.overflow = true;
            return nullptr;
        }
        a.used = off + bytes;
        return a.base + off;
    }
    if (g_rec) *g_rec += (bytes + 255) / 256 * 256 + 256;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto &slots = g_scratch[{dev, stream}];
    if ((int)slots.size() <= slot) slots.resize(slot + 1);
    ScratchBuf &b = slots[slot];
    if (bytes == 0) bytes = 1;
    if (b.bytes < bytes) {
        if (b.ptr) {
            // the stream may still use the old buffer
            if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return nullptr;
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        // grow with headroom so alternating sizes do not reallocate
        const size_t want = bytes + bytes / 8;
        if (hipMalloc(&b.ptr, want) != hipSuccess) {
            (void)hipGetLastError();
            b.ptr = nullptr;
            return nullptr;
        }
        b.bytes = want;
    }
    return b.ptr;
}

static int release_scratch()
{
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    for (auto it = g_scratch.begin(); it != g_scratch.end();) {
        if (it->first.first == dev) {
            MDG_HIP_CHECK(hipStreamSynchronize((hipStream_t)it->first.second));
            for (auto &b : it->second)
                if (b.ptr) (void)hipFree(b.ptr);
            it = g_scratch.erase(it);
        } else {
            ++it;
        }
    }
    return MDG_SUCCESS;
}

static int ensure_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
        return MDG_ERROR_HIP;
    }
    return MDG_SUCCESS;
}

int check_gemm_args(int L, int m, int n, int k, const void *A, long long lda, const void *B, long long ldb,
                           const void *C, long long ldc)
{
    MDG_REQUIRE(valid_limbs(L), MDG_ERROR_INVALID_VALUE, "unsupported limb count L=%d (expected 2, 4, 8 or 16)", L);
    MDG_REQUIRE(m >= 0 && n >= 0 && k >= 0, MDG_ERROR_INVALID_VALUE, "negative dimension (m=%d n=%d k=%d)", m, n, k);
    MDG_REQUIRE(lda >= k && ldb >= n && ldc >= n, MDG_ERROR_INVALID_VALUE,
                "leading dimension too small (lda=%lld >= k=%d, ldb=%lld >= n=%d, ldc=%lld >= n=%d)", lda, k, ldb, n,
                ldc, n);
    MDG_REQUIRE((m == 0 || n == 0) || C, MDG_ERROR_INVALID_VALUE, "C is NULL");
    MDG_REQUIRE((m == 0 || n == 0 || k == 0) || (A && B), MDG_ERROR_INVALID_VALUE, "A or B is NULL");
    return MDG_SUCCESS;
}

__global__ void zero_entries_kernel(double *C, int m, int n, long long ldc, int L)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)m * n * L) return;
    const long long r = idx / ((long long)n * L);
    const long long rem = idx - r * n * L;
    C[r * ldc * L + rem] = 0.0;
}

static int zero_c(double *C, int m, int n, long long ldc, int L, hipStream_t s)
{
    const long long nel = (long long)m * n * L;
    if (nel == 0) return MDG_SUCCESS;
    hipLaunchKernelGGL(zero_entries_kernel, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, s, C, m, n, ldc, L);
    MDG_LAUNCH_CHECK("zero_entries_kernel");
    return MDG_SUCCESS;
}

}  // namespace mdg

using namespace mdg;

extern "C" {

const char *mdg_last_error(void) { return get_error(); }

int mdg_version(void) { return 100; }

int mdg_init(void)
{
    int rc = ensure_device();
    if (rc) return rc;
    MDG_HIP_CHECK(hipFree(nullptr));
    void *ws = nullptr;   // mdg_gemm_noalloc's workspace address on this device
    return noalloc_workspace(&ws);
}

int mdg_release_workspace(void)
{
    int rc = release_scratch();
    return rc ? rc : release_batch_plans();
}

int mdg_set_batch_plan_cache(int max_plans, long long max_bytes)
{
    return set_batch_plan_cache(max_plans, max_bytes);
}

int mdg_batched_last_gemm_ms(void *stream, float *ms)
{
    MDG_REQUIRE(ms, MDG_ERROR_INVALID_VALUE, "NULL ms");
    int rc = ensure_device();
    return rc ? rc : batched_last_gemm_ms(stream, ms);
}

int mdg_batched_last_flags(void *stream, int count, int *flags)
{
    MDG_REQUIRE(count >= 0 && (count == 0 || flags), MDG_ERROR_INVALID_VALUE, "bad count or NULL flags");
    int rc = ensure_device();
    return rc ? rc : batched_last_flags(stream, count, flags);
}

int mdg_set_int8_backend(int backend)
{
    MDG_REQUIRE(backend >= 0 && backend <= 4, MDG_ERROR_INVALID_VALUE, "unknown INT8 backend %d", backend);
    g_int8_backend.store(backend);
    return MDG_SUCCESS;
}

int mdg_set_profiling(int on)
{
    g_profiling.store(on ? 1 : 0);
    return MDG_SUCCESS;
}

int mdg_vector_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                    double *C, long long ldc, int twoprod, void *stream)
{
    int rc = check_gemm_args(L, m, n, k, A, lda, B, ldb, C, ldc);
    if (rc) return rc;
    MDG_REQUIRE(twoprod == MDG_TWOPROD_FMA || twoprod == MDG_TWOPROD_DEKKER, MDG_ERROR_INVALID_VALUE,
                "unknown two_prod mode %d", twoprod);
    if ((rc = ensure_device())) return rc;
    if (m == 0 || n == 0) return MDG_SUCCESS;
    if (k == 0) return zero_c(C, m, n, ldc, L, as_stream(stream));
    return vector_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, twoprod, as_stream(stream));
}

int mdg_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb, double *C,
             long long ldc, int method, int target_bits, mdg_gemm_info *info, void *stream)
{
    int rc = check_gemm_args(L, m, n, k, A, lda, B, ldb, C, ldc);
    if (rc) return rc;
    MDG_REQUIRE(method >= MDG_METHOD_VECTOR && method <= MDG_METHOD_F64_RESIDUE, MDG_ERROR_INVALID_VALUE,
                "unknown method %d", method);
    if ((rc = ensure_device())) return rc;
    if (info) {
        *info = mdg_gemm_info{};
        info->method = method;
    }
    if (m == 0 || n == 0) return MDG_SUCCESS;
    if (k == 0) return zero_c(C, m, n, ldc, L, as_stream(stream));
    if (method == MDG_METHOD_VECTOR) {
        StageTimer tm(as_stream(stream), profiling_on() && info != nullptr);
        tm.mark(0);
        tm.mark(1);
        tm.mark(2);
        rc = vector_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, MDG_TWOPROD_FMA, as_stream(stream));
        if (rc) return rc;
        tm.mark(3);
        tm.mark(4);
        if (tm.on) {
            float st[4];
            MDG_HIP_CHECK(tm.elapsed(st));
            for (int i = 0; i < 4; ++i) info->stage_ms[i] = st[i];
        }
        return MDG_SUCCESS;
    }
    return split_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, method, target_bits, -1, info, as_stream(stream));
}

int mdg_gemm_batched(int count, const mdg_problem *problems, int method, int target_bits, int spread_bits,
                     int persistent, void *stream)
{
    MDG_REQUIRE(count >= 0, MDG_ERROR_INVALID_VALUE, "negative batch count");
    MDG_REQUIRE(count == 0 || problems, MDG_ERROR_INVALID_VALUE, "problems is NULL");
    for (int i = 0; i < count; ++i) {
        const mdg_problem &p = problems[i];
        int rc = check_gemm_args(p.L, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc);
        if (rc) {
            std::string msg = std::string("problem ") + std::to_string(i) + ": " + get_error();
            set_error("%s", msg.c_str());
            return rc;
        }
    }
    MDG_REQUIRE(method >= MDG_METHOD_VECTOR && method <= MDG_METHOD_F64_RESIDUE, MDG_ERROR_INVALID_VALUE,
                "unknown method %d", method);
    int rc = ensure_device();
    if (rc) return rc;
    if (count == 0) return MDG_SUCCESS;
    return batched_gemm(count, problems, method, target_bits, spread_bits, persistent, as_stream(stream));
}

int mdg_split4pd(const double *x, long long count, int p, double *out, void *stream)
{
    MDG_REQUIRE(p >= 1 && count >= 0 && count % p == 0, MDG_ERROR_INVALID_VALUE,
                "count=%lld must be a non-negative multiple of p=%d", count, p);
    MDG_REQUIRE(count == 0 || (x && out), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return ref_split(4, x, count, p, out, as_stream(stream));
}

int mdg_split8pd(const double *x, long long count, int p, double *out, void *stream)
{
    MDG_REQUIRE(p >= 1 && count >= 0 && count % p == 0, MDG_ERROR_INVALID_VALUE,
                "count=%lld must be a non-negative multiple of p=%d", count, p);
    MDG_REQUIRE(count == 0 || (x && out), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return ref_split(8, x, count, p, out, as_stream(stream));
}

int mdg_mixsplit2(const double *x, long long count, double *out, void *stream)
{
    MDG_REQUIRE(count >= 0 && count % 2 == 0, MDG_ERROR_INVALID_VALUE, "count=%lld must be even", count);
    MDG_REQUIRE(count == 0 || (x && out), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return ref_split(12, x, count, 2, out, as_stream(stream));
}

int mdg_bigB2(const double *Bq, int n, int pp, double *out, void *stream)
{
    MDG_REQUIRE(n >= 0 && pp >= 1, MDG_ERROR_INVALID_VALUE, "bad n=%d or pp=%d", n, pp);
    MDG_REQUIRE(n == 0 || (Bq && out), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return bigB2(Bq, n, pp, out, as_stream(stream));
}

int mdg_matmulhost(int p, int n, const double *A, const double *B, double *C, void *stream)
{
    MDG_REQUIRE(n >= 0 && p >= 1, MDG_ERROR_INVALID_VALUE, "bad n=%d or p=%d", n, p);
    MDG_REQUIRE(n == 0 || (A && B && C), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return matmulhost(A, B, n, p, C, as_stream(stream));
}

int mdg_renormhost(double *A, int n, int p, int big, void *stream)
{
    MDG_REQUIRE(n >= 0 && p >= 1, MDG_ERROR_INVALID_VALUE, "bad n=%d or p=%d", n, p);
    MDG_REQUIRE(big == 0 || big == 1, MDG_ERROR_INVALID_VALUE, "big must be 0 or 1 (got %d)", big);
    MDG_REQUIRE(n == 0 || A, MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return renormhost(A, n, p, big, as_stream(stream));
}

int mdg_dgemm_acc(int M, int N, int K, const double *A, const double *B, double *C, void *stream)
{
    MDG_REQUIRE(M >= 0 && N >= 0 && K >= 0, MDG_ERROR_INVALID_VALUE, "negative dimension");
    MDG_REQUIRE(M == 0 || N == 0 || K == 0 || (A && B && C), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    if (M == 0 || N == 0 || K == 0) return MDG_SUCCESS;
    return dgemm_acc(M, N, K, A, B, C, as_stream(stream));
}

int mdg_conv_gemm(int m, int n, int k, int pp, const double *Aq, const double *Bq, double *C1q, void *stream)
{
    MDG_REQUIRE(m >= 0 && n >= 0 && k >= 1 && pp >= 1, MDG_ERROR_INVALID_VALUE, "bad shape m=%d n=%d k=%d pp=%d", m,
                n, k, pp);
    MDG_REQUIRE(pp <= 65535, MDG_ERROR_NOT_SUPPORTED, "pp=%d too large", pp);
    MDG_REQUIRE(m == 0 || n == 0 || (Aq && Bq && C1q), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    if (m == 0 || n == 0) return MDG_SUCCESS;
    return conv_gemm(m, n, k, pp, Aq, Bq, C1q, as_stream(stream));
}

int mdg_squeeze(int L, long long nent, int pp, const double *x, double *y, void *stream)
{
    MDG_REQUIRE(valid_limbs(L), MDG_ERROR_INVALID_VALUE, "unsupported limb count L=%d", L);
    MDG_REQUIRE(nent >= 0 && pp >= 1, MDG_ERROR_INVALID_VALUE, "bad nent=%lld or pp=%d", nent, pp);
    MDG_REQUIRE(nent == 0 || (x && y), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return squeeze(L, nent, pp, x, y, as_stream(stream));
}

int mdg_elementwise(int op, int L, long long n, const double *a, const double *b, double *c, int twoprod,
                    void *stream)
{
    MDG_REQUIRE(op >= MDG_OP_ADD && op <= MDG_OP_NEG, MDG_ERROR_INVALID_VALUE, "unsupported elementwise op %d", op);
    MDG_REQUIRE(valid_limbs(L), MDG_ERROR_INVALID_VALUE, "unsupported limb count L=%d", L);
    MDG_REQUIRE(n >= 0, MDG_ERROR_INVALID_VALUE, "negative entry count %lld", n);
    MDG_REQUIRE(twoprod == MDG_TWOPROD_FMA || twoprod == MDG_TWOPROD_DEKKER, MDG_ERROR_INVALID_VALUE,
                "bad twoprod %d", twoprod);
    const bool needs_b = elementwise_binary(op) || elementwise_takes_double(op);
    MDG_REQUIRE(n == 0 || (a && c && (b || !needs_b)), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    MDG_REQUIRE(n <= (1ll << 40), MDG_ERROR_INVALID_VALUE, "entry count %lld too large", n);
    int rc = ensure_device();
    if (rc) return rc;
    if (n == 0) return MDG_SUCCESS;
    return elementwise(op, L, n, a, b, c, twoprod, as_stream(stream));
}

static int check_layout_args(int L, int rows, int cols, const void *x, long long ldx, const void *y)
{
    MDG_REQUIRE(valid_limbs(L), MDG_ERROR_INVALID_VALUE, "unsupported limb count L=%d", L);
    MDG_REQUIRE(rows >= 0 && cols >= 0, MDG_ERROR_INVALID_VALUE, "bad shape %d x %d", rows, cols);
    MDG_REQUIRE(rows == 0 || cols == 0 || ldx >= cols, MDG_ERROR_INVALID_VALUE, "row stride %lld < cols %d", ldx,
                cols);
    MDG_REQUIRE(rows == 0 || cols == 0 || (x && y), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    MDG_REQUIRE((long long)((rows + 15) / 16) * ((cols + 15) / 16) < (1ll << 31) &&
                    (long long)rows * ((cols + 255) / 256) < (1ll << 31),
                MDG_ERROR_INVALID_VALUE, "matrix %d x %d too large", rows, cols);
    return MDG_SUCCESS;
}

int mdg_to_planar(int L, int rows, int cols, const double *x, long long ldx, double *planar, void *stream)
{
    int rc = check_layout_args(L, rows, cols, x, ldx, planar);
    if (rc || (rc = ensure_device())) return rc;
    if (rows == 0 || cols == 0) return MDG_SUCCESS;
    return to_planar(L, rows, cols, x, ldx, planar, as_stream(stream));
}

int mdg_from_planar(int L, int rows, int cols, const double *planar, double *x, long long ldx, void *stream)
{
    int rc = check_layout_args(L, rows, cols, x, ldx, planar);
    if (rc || (rc = ensure_device())) return rc;
    if (rows == 0 || cols == 0) return MDG_SUCCESS;
    return from_planar(L, rows, cols, planar, x, ldx, as_stream(stream));
}

int mdg_transpose(int L, int rows, int cols, const double *x, long long ldx, double *y, long long ldy, void *stream)
{
    int rc = check_layout_args(L, rows, cols, x, ldx, y);
    if (rc) return rc;
    MDG_REQUIRE(rows == 0 || cols == 0 || ldy >= rows, MDG_ERROR_INVALID_VALUE, "row stride %lld < rows %d", ldy,
                rows);
    MDG_REQUIRE(rows == 0 || cols == 0 || x != y, MDG_ERROR_INVALID_VALUE, "in-place transpose is not supported");
    if ((rc = ensure_device())) return rc;
    if (rows == 0 || cols == 0) return MDG_SUCCESS;
    return transpose(L, rows, cols, x, ldx, y, ldy, as_stream(stream));
}

int mdg_tc_gemm(int L, int m, int n, int k, const double *A, const double *B, double *C, int split_kind,
                void *stream)
{
    MDG_REQUIRE(valid_limbs(L), MDG_ERROR_INVALID_VALUE, "unsupported limb count L=%d", L);
    MDG_REQUIRE(split_kind == 4 || split_kind == 8 || (split_kind == 12 && L == 2), MDG_ERROR_INVALID_VALUE,
                "split_kind must be 4, 8, or 12 (L == 2 only)");
    MDG_REQUIRE(m >= 0 && n >= 0 && k >= 1, MDG_ERROR_INVALID_VALUE, "bad shape m=%d n=%d k=%d", m, n, k);
    MDG_REQUIRE(m == 0 || n == 0 || (A && B && C), MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    if (m == 0 || n == 0) return MDG_SUCCESS;
    hipStream_t s = as_stream(stream);
    const int per = (split_kind == 12) ? 6 : split_kind;
    const int pp = per * L;
    double *Aq = (double *)scratch(stream, sizeof(double) * (size_t)m * k * pp, 5);
    double *Bq = (double *)scratch(stream, sizeof(double) * (size_t)k * n * pp, 6);
    MDG_REQUIRE(Aq && Bq, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    if ((rc = ref_split(split_kind, A, (long long)m * k * L, L, Aq, s))) return rc;
    if ((rc = ref_split(split_kind, B, (long long)k * n * L, L, Bq, s))) return rc;
    return conv_squeeze(L, m, n, k, pp, Aq, Bq, C, s);
}

int mdg_fill_random(int kind, int L, long long count, unsigned long long seed, int emin, int emax, double *out,
                    void *stream)
{
    MDG_REQUIRE(kind >= 0 && kind <= 2, MDG_ERROR_INVALID_VALUE, "unknown generator kind %d", kind);
    MDG_REQUIRE(L >= 1 && L <= 64, MDG_ERROR_INVALID_VALUE, "bad limb count %d", L);
    MDG_REQUIRE(count >= 0 && emax >= emin, MDG_ERROR_INVALID_VALUE, "bad count or exponent range");
    MDG_REQUIRE(count == 0 || out, MDG_ERROR_INVALID_VALUE, "NULL pointer");
    int rc = ensure_device();
    if (rc) return rc;
    return fill_random(kind, L, count, seed, emin, emax, out, as_stream(stream));
}

}  // extern "C"
