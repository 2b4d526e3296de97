// api_internal.h - host-side state shared by the C-ABI implementation files:
// the context and the error helpers (errors are returned with a message kept
// in the context, never aborted on).
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "internal.h"

// frames of a batch are shaded on this many streams at once (forked from and
// joined back into the context's stream), each with its own workspace
constexpr int SHADE_STREAMS = 8;

struct hr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t side[SHADE_STREAMS] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[SHADE_STREAMS] = {};
    unsigned long long* d_counters = nullptr;
    uint32_t* d_next_group = nullptr;  // work counter of the persistent shadow kernel
    uint32_t cu_count = 256;
    uint32_t resident_blocks = 2048;    // persistent grid: cu_count x tuning.blocks_per_cu
    hr_tuning tuning{8u, 8u, 40960u, 0u};  // 40 GiB reflection workspace at most (1080p at 64 samples per pixel: one slab)
    void* work = nullptr;  // grow-only workspace of the reflection pipeline
    size_t work_bytes = 0;
    hr::ShadowJob* d_jobs = nullptr;  // grow-only device copy of the shadow jobs of a call
    uint32_t jobs_cap = 0;
    // this device's address of the static scratch of hr_trace_shadows_noalloc
    // (null when hipGetSymbolAddress failed: noalloc_err)
    void* noalloc_scratch = nullptr;
    hipError_t noalloc_err = hipSuccess;
    std::string err;
};

namespace hr {

inline hr_status fail(hr_context* ctx, hr_status code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}
inline hr_status hip_fail(hr_context* ctx, hipError_t e, const char* what) {
    hr_status code = (e == hipErrorOutOfMemory) ? HR_E_OOM : HR_E_DEVICE;
    return fail(ctx, code, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace hr

#define HR_HIP(ctx, call)                                          \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return hr::hip_fail(ctx, e_, #call); \
    } while (0)
#define HR_TRY(x)                     \
    do {                              \
        hr_status st_ = (x);          \
        if (st_ != HR_OK) return st_; \
    } while (0)
