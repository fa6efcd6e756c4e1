// mcl_ctx.h — the library context shared by the Monte Carlo and surrogate
// entry points (not part of the ABI; include/mcl.h declares it opaque).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

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
    double waves_per_cluster = 0.0; // wavefronts per fixed-dt cluster: 1, 2, 4, 8 or 0 (chosen per launch)
    std::string err;
};

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
