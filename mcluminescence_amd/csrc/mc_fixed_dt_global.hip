// mc_fixed_dt_global.hip — the fixed-dt kernels in their global-memory form, for
// clusters whose holes (16 B each) and cell starts do not fit one workgroup's
// LDS beside the electron arrays: the same cluster body (mc_fixed_dt.h) with the
// holes and cell starts in a scratch slice per resident workgroup, launched
// persistent so that the slices are reused cluster after cluster.
#include "mc_fixed_dt.h"

namespace mcl {

hipError_t prepare_fixed_dt_global(int max_lds)
{
    hipError_t e = hipSuccess;
    for (int w : {1, 2, 4, 8}) {
        for (bool fw : {false, true}) {
            if (e == hipSuccess)
                e = hipFuncSetAttribute(fixed_dt_fn<true>(w, fw),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        }
    }
    return e;
}

hipError_t launch_fixed_dt_global(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                  hipStream_t stream)
{
    if (a.queue == nullptr || a.gscratch == nullptr) return hipErrorInvalidValue;
    return launch_fixed_dt_t<true>(a, grid, (unsigned)(lds_bytes + fixed_dt_waves_lds(waves)),
                                   waves, stream);
}

hipError_t fixed_dt_global_occupancy(int lds_bytes, int waves, int* blocks_per_cu)
{
    return fixed_dt_occupancy_t<true>((size_t)(lds_bytes + fixed_dt_waves_lds(waves)), waves,
                                      blocks_per_cu);
}

}  // namespace mcl
