// i8gemm.hip -- INT8 x INT8 -> INT32 strided-batched GEMMs: this library's
// MFMA kernel (plain_i8_kernel, i8plain.hpp) by default; the library engines
// (rocBLAS, hipBLASLt, autotuned per shape) when mdg_set_int8_backend picks
// them, as a comparison, or for a shape the kernel does not take.
//
// Layout (column-major view, the convention of both libraries): for batch b
//   D_b (n x m, leading dimension ldd) = Bt_b^T (n x depth) * A_b (depth x m)
// with Bt_b[c * ldb + kk] and A_b[r * lda + kk] -- both operands contiguous
// along the depth, i.e. D_b is the row-major m x n product of A_b (m x depth)
// and Bt_b^T.  Sums are exact INT32 as long as depth * 2^14 < 2^31.
//
// The first call of a shape bucket (power-of-two buckets of batch, m, n and
// depth) times the first algorithms hipBLASLt proposes for it and rocBLAS's
// default, twice each, and keeps the fastest; all of them compute the same
// exact integers.  A kept hipBLASLt algorithm that rejects another shape of
// its bucket falls back to rocBLAS for that call.
#include <hipblaslt/hipblaslt.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "i8plain.hpp"
#include "split_common.hpp"

namespace mdg {

// the hand-written engine takes depths in whole 128-byte k-blocks and
// 16-byte aligned operand rows
static bool plain_fits(int batch, int depth, const int8_t *Bt, long long ldb, long long sb, const int8_t *A,
                       long long lda, long long sa)
{
    auto al16 = [](long long v) { return (v & 15) == 0; };
    return batch <= 65535 && depth > 0 && depth % PLN_BK == 0 && al16(ldb) && al16(lda) && al16(sb) && al16(sa) &&
           al16((long long)(uintptr_t)Bt) && al16((long long)(uintptr_t)A);
}

static int run_plain(int batch, int m, int n, int depth, const int8_t *Bt, long long ldb, long long sb,
                     const int8_t *A, long long lda, long long sa, int *D, long long ldd, long long sd, hipStream_t s)
{
    const int tiles_m = (m + PLN_TM - 1) / PLN_TM, tiles_n = (n + PLN_TM - 1) / PLN_TM;
    const dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)batch);
    // rows past m (n) of a 256-row tile reread row m-1 (n-1): always readable
    hipLaunchKernelGGL(plain_i8_kernel<0>, grid, dim3(256), 0, s, A, lda, sa, m, Bt, ldb, sb, n, m, n, depth, D, ldd,
                       sd, tiles_m, tiles_n);
    MDG_LAUNCH_CHECK("plain_i8_kernel");
    return MDG_SUCCESS;
}

namespace {

constexpr size_t LT_WORKSPACE = 32ull << 20;

struct LtHandle {
    hipblasLtHandle_t h = nullptr;
    std::mutex mu;
};

std::mutex g_lt_mu;
std::map<int, LtHandle *> g_lt;

LtHandle *lt_handle()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_lt_mu);
    auto it = g_lt.find(dev);
    if (it != g_lt.end()) return it->second;
    LtHandle *l = new LtHandle();
    if (hipblasLtCreate(&l->h) != HIPBLAS_STATUS_SUCCESS) {
        delete l;
        return nullptr;
    }
    g_lt[dev] = l;
    return l;
}

typedef std::tuple<int, int, int, int, int, int> ShapeKey;

constexpr int LT_CANDIDATES = 8;

int lg2_ceil(long long v)
{
    int b = 0;
    while ((1ll << b) < v) ++b;
    return b;
}

struct Choice {
    int engine = 0;   // 1 rocBLAS, 3 hipBLASLt
    hipblasLtMatmulAlgo_t algo{};
};

std::mutex g_choice_mu;
std::map<ShapeKey, Choice> g_choice;

struct LtDescs {
    hipblasLtMatmulDesc_t desc = nullptr;
    hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
    bool ok = false;
    LtDescs(int batch, int m, int n, int depth, long long ldb, long long sb, long long lda, long long sa,
            long long ldd, long long sd)
    {
        if (hipblasLtMatmulDescCreate(&desc, HIPBLAS_COMPUTE_32I, HIP_R_32I) != HIPBLAS_STATUS_SUCCESS) return;
        hipblasOperation_t ta = HIPBLAS_OP_T, tb = HIPBLAS_OP_N;
        if (hipblasLtMatmulDescSetAttribute(desc, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof(ta)) ||
            hipblasLtMatmulDescSetAttribute(desc, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof(tb)))
            return;
        if (hipblasLtMatrixLayoutCreate(&la, HIP_R_8I, depth, n, ldb) ||
            hipblasLtMatrixLayoutCreate(&lb, HIP_R_8I, depth, m, lda) ||
            hipblasLtMatrixLayoutCreate(&lc, HIP_R_32I, n, m, ldd))
            return;
        const int32_t b32 = batch;
        const int64_t strides[3] = {sb, sa, sd};
        hipblasLtMatrixLayout_t ls[3] = {la, lb, lc};
        for (int i = 0; i < 3; ++i)
            if (hipblasLtMatrixLayoutSetAttribute(ls[i], HIPBLASLT_MATRIX_LAYOUT_BATCH_COUNT, &b32, sizeof(b32)) ||
                hipblasLtMatrixLayoutSetAttribute(ls[i], HIPBLASLT_MATRIX_LAYOUT_STRIDED_BATCH_OFFSET, &strides[i],
                                                  sizeof(int64_t)))
                return;
        ok = true;
    }
    ~LtDescs()
    {
        if (la) (void)hipblasLtMatrixLayoutDestroy(la);
        if (lb) (void)hipblasLtMatrixLayoutDestroy(lb);
        if (lc) (void)hipblasLtMatrixLayoutDestroy(lc);
        if (desc) (void)hipblasLtMatmulDescDestroy(desc);
    }
};

}  // namespace

int i8_gemm_tn(int batch, int m, int n, int depth, const int8_t *Bt, long long ldb, long long sb, const int8_t *A,
               long long lda, long long sa, int *D, long long ldd, long long sd, hipStream_t s, int *engine_used)
{
    if (batch <= 0 || m <= 0 || n <= 0) return MDG_SUCCESS;
    // 1 rocBLAS, 3 hipBLASLt (comparison engines); otherwise the MFMA kernel
    const int forced = int8_backend();
    if (forced != 1 && forced != 3 && plain_fits(batch, depth, Bt, ldb, sb, A, lda, sa)) {
        if (engine_used) *engine_used = 2;
        return run_plain(batch, m, n, depth, Bt, ldb, sb, A, lda, sa, D, ldd, sd, s);
    }
    const int32_t alpha = 1, beta = 0;
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    const ShapeKey key{dev, forced, lg2_ceil(batch), lg2_ceil(m), lg2_ceil(n), lg2_ceil(depth)};

    BlasHandle *bh = blas_handle();
    MDG_REQUIRE(bh, MDG_ERROR_HIP, "rocblas_create_handle failed");
    LtHandle *lh = lt_handle();

    auto run_rocblas = [&]() -> int {
        std::lock_guard<std::mutex> lock(bh->mu);
        if (rocblas_set_stream(bh->h, s) != rocblas_status_success) {
            set_error("rocblas_set_stream failed");
            return MDG_ERROR_HIP;
        }
        rocblas_status st = rocblas_gemm_strided_batched_ex(
            bh->h, rocblas_operation_transpose, rocblas_operation_none, n, m, depth, &alpha, Bt,
            rocblas_datatype_i8_r, ldb, sb, A, rocblas_datatype_i8_r, lda, sa, &beta, D, rocblas_datatype_i32_r, ldd,
            sd, D, rocblas_datatype_i32_r, ldd, sd, batch, rocblas_datatype_i32_r, rocblas_gemm_algo_standard, 0, 0);
        if (st != rocblas_status_success) {
            set_error("rocblas_gemm_strided_batched_ex (int8) failed with status %d", (int)st);
            return MDG_ERROR_HIP;
        }
        return MDG_SUCCESS;
    };
    void *ws = nullptr;
    auto run_lt = [&](const LtDescs &d, const hipblasLtMatmulAlgo_t &algo) -> int {
        std::lock_guard<std::mutex> lock(lh->mu);
        hipblasStatus_t st = hipblasLtMatmul(lh->h, d.desc, &alpha, Bt, d.la, A, d.lb, &beta, D, d.lc, D, d.lc, &algo,
                                             ws, LT_WORKSPACE, s);
        if (st != HIPBLAS_STATUS_SUCCESS) {
            set_error("hipblasLtMatmul (int8) failed with status %d", (int)st);
            return MDG_ERROR_HIP;
        }
        return MDG_SUCCESS;
    };

    Choice ch;
    bool known = false;
    {
        std::lock_guard<std::mutex> lock(g_choice_mu);
        auto it = g_choice.find(key);
        if (it != g_choice.end()) {
            ch = it->second;
            known = true;
        }
    }
    if (known && ch.engine == 1) {
        if (engine_used) *engine_used = 1;
        return run_rocblas();
    }
    if (forced == 1 || !lh) {
        if (engine_used) *engine_used = 1;
        return run_rocblas();
    }
    ws = scratch(s, LT_WORKSPACE, 5);
    MDG_REQUIRE(ws, MDG_ERROR_OUT_OF_MEMORY, "scratch allocation failed");
    LtDescs d(batch, m, n, depth, ldb, sb, lda, sa, ldd, sd);
    MDG_REQUIRE(d.ok, MDG_ERROR_HIP, "hipBLASLt descriptor set-up failed");
    if (known) {
        if (run_lt(d, ch.algo) == MDG_SUCCESS) {
            if (engine_used) *engine_used = 3;
            return MDG_SUCCESS;
        }
        if (engine_used) *engine_used = 1;
        return run_rocblas();
    }

    // autotune: hipBLASLt's candidates (and rocBLAS unless hipBLASLt is forced)
    hipblasLtMatmulPreference_t pref = nullptr;
    MDG_REQUIRE(hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS, MDG_ERROR_HIP,
                "hipblasLtMatmulPreferenceCreate failed");
    const uint64_t wmax = LT_WORKSPACE;
    (void)hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wmax, sizeof(wmax));
    std::vector<hipblasLtMatmulHeuristicResult_t> res(LT_CANDIDATES);
    int got = 0;
    {
        std::lock_guard<std::mutex> lock(lh->mu);
        if (hipblasLtMatmulAlgoGetHeuristic(lh->h, d.desc, d.la, d.lb, d.lc, d.lc, pref, (int)res.size(), res.data(),
                                            &got) != HIPBLAS_STATUS_SUCCESS)
            got = 0;
    }
    (void)hipblasLtMatmulPreferenceDestroy(pref);
    hipEvent_t e0, e1;
    MDG_HIP_CHECK(hipEventCreate(&e0));
    MDG_HIP_CHECK(hipEventCreate(&e1));
    auto time_one = [&](auto fn, float &best) -> int {
        best = 1e30f;
        for (int round = 0; round < 2; ++round) {
            MDG_HIP_CHECK(hipEventRecord(e0, s));
            int rc = fn();
            if (rc) return rc;
            MDG_HIP_CHECK(hipEventRecord(e1, s));
            MDG_HIP_CHECK(hipEventSynchronize(e1));
            float t = 0.f;
            MDG_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
            best = std::min(best, t);
        }
        return MDG_SUCCESS;
    };
    Choice best;
    float best_ms = 1e30f;
    int rc = MDG_SUCCESS;
    for (int i = 0; i < got && rc == MDG_SUCCESS; ++i) {
        if (res[i].state != HIPBLAS_STATUS_SUCCESS || res[i].workspaceSize > LT_WORKSPACE) continue;
        float t = 0.f;
        const int r = time_one([&] { return run_lt(d, res[i].algo); }, t);
        if (r == MDG_SUCCESS && t < best_ms) {
            best_ms = t;
            best.engine = 3;
            best.algo = res[i].algo;
        }
    }
    if (forced != 3) {
        float t = 0.f;
        rc = time_one(run_rocblas, t);
        if (rc == MDG_SUCCESS && t < best_ms) {
            best_ms = t;
            best.engine = 1;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    MDG_REQUIRE(best.engine != 0, MDG_ERROR_HIP, "no INT8 GEMM engine ran for %d x %d x %d (batch %d)", m, n, depth,
                batch);
    {
        std::lock_guard<std::mutex> lock(g_choice_mu);
        g_choice[key] = best;
    }
    // every engine writes the same exact D, but a candidate that failed may
    // have been the last to touch it: the winner computes it once more
    if (engine_used) *engine_used = best.engine;
    return best.engine == 1 ? run_rocblas() : run_lt(d, best.algo);
}

}  // namespace mdg
