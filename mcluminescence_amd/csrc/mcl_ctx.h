// mcl_ctx.h — the library context shared by the Monte Carlo and surrogate
// entry points (not part of the ABI; include/mcl.h declares it opaque).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/mcl.h"

namespace mcl {

struct Ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int max_lds = 0;        // dynamic LDS bytes one workgroup may use
    int n_cu = 0;
    double holes_per_cell = 2.0;   // hole-grid density target (tunable)
    double persist_frac = 1.0;     // persistent workgroups per class / device capacity
    double wave_search_max = 32.0; // queue length searched by the whole wave
    double lookahead_max = 8.0;    // live electrons up to which quiet steps are skipped
    double quiet_window = 1.0;     // 1: after a step with no candidate, the rest of its
                                   // Philox block is tested at once
    double waves_per_cluster = 0.0; // wavefronts per fixed-dt cluster: 1, 2, 4, 8 or 0 (chosen per launch)
    double mlp_fused = 1.0;        // surrogate inference as one kernel (1) or per layer (0)
    double global_memory = 0.0;    // fixed-dt clusters in the global-memory form: 0 when
                                   // they exceed LDS, 1 always
    double fixed_thresholds = 1.0; // 1: launches with an isothermal one-channel run take the
                                   // kernels with the fixed-threshold path; 0: never
    std::string err;
    // timing events of mcl_run_fixed_dt_noalloc, made once by mcl_create
    hipEvent_t na_ev0 = nullptr, na_ev1 = nullptr;
    // Device blocks of finished runs, kept for later runs (hipMalloc and hipFree,
    // which waits for the device, cost more than a small run's kernels).
    std::mutex pool_mu;
    std::vector<std::pair<size_t, void*>> pool_free;   // (bytes, block), idle
    std::unordered_map<void*, size_t> pool_live;        // handed out: block -> bytes
    size_t pool_cached = 0;                             // bytes in pool_free
};

// A device block of at least `bytes` from the context's pool, else hipMalloc.
hipError_t pool_alloc(Ctx* c, void** p, size_t bytes);
// Returns a block to the pool.  idle: every kernel and copy that used it has
// finished (the caller synchronised the stream it ran on); otherwise the device
// is synchronised first, as hipFree would.
void pool_release(Ctx* c, void* p, bool idle);
// Frees the idle blocks (at context destruction).
void pool_clear(Ctx* c);

// Record a formatted message as the context's last error and return code.
inline int fail(Ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

}  // namespace mcl

#define HIP_TRY(ctx, expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return mcl::fail((ctx), e_ == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE,     \
                             "%s: %s", #expr, hipGetErrorString(e_));                         \
    } while (0)
