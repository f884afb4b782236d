// noalloc.hip -- mdg_gemm_noalloc: the split-product / vector GEMM of
// mdg_gemm with all of its workspace in one static device array, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  No counterpart in the
// reference.
#include <map>
#include <mutex>

#include "common.hpp"

// The workspace: NOALLOC_BYTES of device memory that every GPU the library
// runs on holds from module load until the process ends.  Host code never
// uses this name as an address; it uses the per-device pointer that
// hipGetSymbolAddress returns.
static constexpr size_t NOALLOC_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace mdg {

int check_gemm_args(int L, int m, int n, int k, const void *A, long long lda, const void *B, long long ldb,
                    const void *C, long long ldc);

static std::mutex g_ws_mu;
static std::map<int, void *> g_ws;   // device -> address of `scratch` there

// the workspace address on the current device, looked up once per device
int noalloc_workspace(void **ptr)
{
    int dev = 0;
    MDG_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ws_mu);
    auto it = g_ws.find(dev);
    if (it != g_ws.end()) {
        *ptr = it->second;
        return MDG_SUCCESS;
    }
    void *p = nullptr;
    MDG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));   // the device array, not mdg::scratch()
    MDG_REQUIRE(p != nullptr, MDG_ERROR_HIP, "hipGetSymbolAddress returned no address for the workspace");
    g_ws[dev] = p;
    *ptr = p;
    return MDG_SUCCESS;
}

}  // namespace mdg

using namespace mdg;

extern "C" {

size_t mdg_noalloc_workspace_bytes(void) { return NOALLOC_BYTES; }

int mdg_gemm_noalloc(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                     double *C, long long ldc, int method, int target_bits, mdg_gemm_info *info, void *stream)
{
    int rc = check_gemm_args(L, m, n, k, A, lda, B, ldb, C, ldc);
    if (rc) return rc;
    MDG_REQUIRE(method == MDG_METHOD_VECTOR || method == MDG_METHOD_F64_SPLIT || method == MDG_METHOD_INT8_OZAKI ||
                    method == MDG_METHOD_BF16_OZAKI,
                MDG_ERROR_NOT_SUPPORTED,
                "mdg_gemm_noalloc runs the vector, FP64-split, INT8-slice and BF16-slice methods (method %d)", method);
    if (m == 0 || n == 0 || k == 0 || method == MDG_METHOD_VECTOR)
        return mdg_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, method, target_bits, info, stream);
    void *ws = nullptr;
    rc = noalloc_workspace(&ws);
    if (rc) return rc;
    Arena arena;
    arena_begin(ws, NOALLOC_BYTES, &arena);
    arena.noalloc = true;
    rc = mdg_gemm(L, m, n, k, A, lda, B, ldb, C, ldc, method, target_bits, info, stream);
    size_t need = 0;
    const bool over = arena_overflow(&need);
    arena_end();
    if (over) {
        set_error("mdg_gemm_noalloc: this %d x %d x %d problem (L = %d) needs at least %zu bytes of workspace, more "
                  "than the %zu-byte static workspace; call mdg_gemm instead",
                  m, n, k, L, need, NOALLOC_BYTES);
        return MDG_ERROR_OUT_OF_MEMORY;
    }
    return rc;
}

}  // extern "C"
