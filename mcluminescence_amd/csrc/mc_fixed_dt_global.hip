// mc_fixed_dt_global.hip — the fixed-dt kernels in their global-memory form, for
// clusters whose holes (16 B each) and cell starts do not fit one workgroup's
// LDS beside the electron arrays: the same cluster body (mc_fixed_dt.h) with the
// holes and cell starts in a scratch slice per resident workgroup, launched
// persistent so that the slices are reused cluster after cluster.
#include "mc_fixed_dt.h"

namespace mcl {

hipError_t prepare_fixed_dt_global(int max_lds)
{
    // every build may take the whole of a workgroup's LDS
    hipError_t e = prepare_fixed_dt_t<true, false>(max_lds, max_lds);
    if (e == hipSuccess) e = prepare_fixed_dt_global_window(max_lds);
    return e;
}

hipError_t launch_fixed_dt_global(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                  hipStream_t stream)
{
    if (a.queue == nullptr || a.gscratch == nullptr) return hipErrorInvalidValue;
    if (a.quiet_window != 0) return launch_fixed_dt_global_window(a, grid, lds_bytes, waves, stream);
    return launch_fixed_dt_t<true, false>(a, grid, (unsigned)(lds_bytes + fixed_dt_waves_lds(waves)),
                                          waves, stream);
}

hipError_t fixed_dt_global_occupancy(int lds_bytes, int waves, int* blocks_per_cu)
{
    int b0 = 0, b1 = 0;
    hipError_t e = fixed_dt_occupancy_t<true, false>((size_t)(lds_bytes + fixed_dt_waves_lds(waves)),
                                                     waves, &b0);
    if (e == hipSuccess) e = fixed_dt_global_window_occupancy(lds_bytes, waves, &b1);
    *blocks_per_cu = b0 < b1 ? b0 : b1;
    return e;
}

}  // namespace mcl
