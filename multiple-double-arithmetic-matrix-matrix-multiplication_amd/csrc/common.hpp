// common.hpp -- status plumbing shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mdgemm.h"

namespace mdg {

// thread-local text of the last failure (mdg_last_error)
void set_error(const char *fmt, ...);
const char *get_error();

struct Status {
    int code;
};

#define MDG_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            ::mdg::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                  \
            return MDG_ERROR_HIP;                                                  \
        }                                                                          \
    } while (0)

#define MDG_REQUIRE(cond, code, ...)      \
    do {                                  \
        if (!(cond)) {                    \
            ::mdg::set_error(__VA_ARGS__); \
            return (code);                \
        }                                 \
    } while (0)

// check the launch that was just enqueued
#define MDG_LAUNCH_CHECK(name)                                                            \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess) {                                                           \
            ::mdg::set_error("launch of %s failed: %s", name, hipGetErrorString(_e));     \
            return MDG_ERROR_HIP;                                                         \
        }                                                                                 \
    } while (0)

inline bool valid_limbs(int L) { return L == 2 || L == 4 || L == 8 || L == 16; }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Device scratch, cached per (device, stream) and grown on demand.  Calls
// that share a stream are ordered by it; different streams get different
// buffers, so concurrent streams never alias scratch.
void *scratch(void *stream, size_t bytes, int slot);

// Workspace arena of mdg_gemm_noalloc (api.hip): between arena_begin and
// arena_end, scratch() on this thread returns pieces of [base, base + cap)
// and allocates nothing; arena_overflow reports the bytes a call needed
// when they did not fit.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0, need = 0;
    bool overflow = false;
    bool noalloc = false;   // installed by mdg_gemm_noalloc: hand-written kernels only
};
void arena_begin(void *base, size_t cap, Arena *a);
void arena_end();
void arena_reset();   // start carving from the base again (the next problem of a loop)
bool arena_active();
bool noalloc_active();
bool arena_overflow(size_t *need);
// record the sizes of scratch() requests made on this thread (NULL: stop)
void scratch_record(size_t *counter);

// batched.hip: per-problem flags of a fixed-spread call (split_gemm.hip)
void set_guard_flag(int *flag);

// mdg_set_profiling / mdg_set_int8_backend state
bool profiling_on();
int int8_backend();

// HIP events at stage boundaries of one call, on the call's stream
struct StageTimer {
    bool on;
    hipStream_t s;
    hipEvent_t ev[5] = {};
    StageTimer(hipStream_t stream, bool enabled) : on(enabled), s(stream)
    {
        if (on)
            for (auto &e : ev)
                if (hipEventCreate(&e) != hipSuccess) on = false;
    }
    void mark(int i)
    {
        if (on) (void)hipEventRecord(ev[i], s);
    }
    // elapsed ms between consecutive marks 0..4 (waits for the last mark)
    hipError_t elapsed(float *ms)
    {
        hipError_t e = hipEventSynchronize(ev[4]);
        for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
        return e;
    }
    ~StageTimer()
    {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// counter-based generator used by mdg_fill_random and the split kernels' tests
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace mdg
