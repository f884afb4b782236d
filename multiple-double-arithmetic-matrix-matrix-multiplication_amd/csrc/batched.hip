// batched.hip -- many independent multi-double GEMMs of uneven shape.
//
// persistent = 0: a loop of per-problem launches on the caller's stream
//                 (the baseline a caller of the reference would write around
//                 matmulTCnt, C++/TCcuda.cu:248-302).
// persistent = 1: the same work recorded once into a hipGraph keyed by the
//                 batch's shape signature and replayed on later calls, so the
//                 per-launch CPU cost is paid once.
// Slice counts come from the caller's spread allowance, never from a
// device->host read, so nothing in the batch synchronises the stream.
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace mdg {

int vector_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb,
                double *C, long long ldc, int twoprod, hipStream_t s);
int split_gemm(int L, int m, int n, int k, const double *A, long long lda, const double *B, long long ldb, double *C,
               long long ldc, int method, int target_bits, int fixed_spread, mdg_gemm_info *info, hipStream_t s);

static int run_loop(int count, const mdg_problem *pr, int method, int target_bits, int spread, hipStream_t s)
{
    for (int i = 0; i < count; ++i) {
        const mdg_problem &p = pr[i];
        if (p.m == 0 || p.n == 0 || p.k == 0) continue;
        int rc = (method == MDG_METHOD_VECTOR)
                     ? vector_gemm(p.L, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc, MDG_TWOPROD_FMA, s)
                     : split_gemm(p.L, p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc, method, target_bits,
                                  spread, nullptr, s);
        if (rc) return rc;
    }
    return MDG_SUCCESS;
}

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
};

static std::mutex g_graph_mu;
static std::map<std::string, GraphEntry> g_graphs;

static std::string signature(int count, const mdg_problem *pr, int method, int target_bits, int spread, hipStream_t s)
{
    std::string key;
    key.reserve(64 + sizeof(mdg_problem) * (size_t)count);
    key.append((const char *)&method, sizeof(method));
    key.append((const char *)&target_bits, sizeof(target_bits));
    key.append((const char *)&spread, sizeof(spread));
    key.append((const char *)&s, sizeof(s));
    key.append((const char *)pr, sizeof(mdg_problem) * (size_t)count);
    return key;
}

int batched_gemm(int count, const mdg_problem *pr, int method, int target_bits, int spread_bits, int persistent,
                 hipStream_t s)
{
    const int spread = spread_bits < 0 ? 16 : spread_bits;
    if (!persistent) return run_loop(count, pr, method, target_bits, spread, s);

    const std::string key = signature(count, pr, method, target_bits, spread, s);
    std::lock_guard<std::mutex> lock(g_graph_mu);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
        // one uncaptured pass sizes every scratch buffer (capture must not allocate)
        int rc = run_loop(count, pr, method, target_bits, spread, s);
        if (rc) return rc;
        MDG_REQUIRE(s != nullptr, MDG_ERROR_NOT_SUPPORTED, "graph replay needs a non-null stream");
        MDG_HIP_CHECK(hipStreamSynchronize(s));
        hipGraph_t graph = nullptr;
        MDG_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = run_loop(count, pr, method, target_bits, spread, s);
        hipError_t ce = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        MDG_HIP_CHECK(ce);
        GraphEntry e;
        hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MDG_HIP_CHECK(ie);
        g_graphs[key] = e;
        return MDG_SUCCESS;   // the uncaptured pass already computed this call
    }
    MDG_HIP_CHECK(hipGraphLaunch(it->second.exec, s));
    return MDG_SUCCESS;
}

}  // namespace mdg
