// ctx_internal.h -- the device context shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <string>

#include "disco_hip.h"

struct disco_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    // cached atom buffers (grow-only)
    double* d_atoms = nullptr;   // x | y | z | ux | uy | uz
    size_t cap_atoms = 0;
    // output stage (grow-only): per-block byte counts and offsets, a device
    // text buffer, and two pinned staging buffers for file writes
    void* d_text_scratch = nullptr;
    size_t cap_text_scratch = 0;
    char* d_text = nullptr;
    size_t cap_text = 0;
    char* h_stage[2] = {nullptr, nullptr};
    // prismatic slab selection (grow-only): loop copy, block counts / offsets, flags
    void* d_slab_scratch = nullptr;
    size_t cap_slab_scratch = 0;
    void* h_slab_stage = nullptr;   // pinned: the loop copy and the slab total
    // optional timing of the text kernels of the last format call (count, scan, write)
    int text_profile = 0;
    hipEvent_t text_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float text_ms[3] = {0.f, 0.f, 0.f};
    // this device's copy of the library's static scratch array (the address
    // hipGetSymbolAddress gives for it), used by disco_hip_plan_run_noalloc
    void* noalloc_scratch = nullptr;
};

namespace disco {

inline int fail(disco_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

// Free the output stage's (writers.hip) and the slab selection's
// (prismatic.hip) buffers; called by disco_hip_close.
void release_output_buffers(disco_ctx* ctx);
void release_prismatic_buffers(disco_ctx* ctx);

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace disco

// Returns from the calling function with DISCO_ERR_HIP (or DISCO_ERR_OOM) and
// records the failing call on the context.
#define DISCO_HIP_TRY(ctx, call)                                                        \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return disco::fail((ctx), _e == hipErrorOutOfMemory ? DISCO_ERR_OOM : DISCO_ERR_HIP, \
                               std::string(#call) + ": " + hipGetErrorString(_e));      \
    } while (0)
