// mc_fixed_dt_global_window.hip — the fixed-dt kernels in their global-memory form,
// the builds with the quiet window (mc_fixed_dt.h), taken by launches with a
// ramped run.  A source of their own, so they compile beside the others.
#include "mc_fixed_dt.h"

namespace mcl {

hipError_t prepare_fixed_dt_global_window(int max_lds)
{
    return prepare_fixed_dt_t<true, true>(max_lds, max_lds);
}

hipError_t launch_fixed_dt_global_window(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                         hipStream_t stream)
{
    return launch_fixed_dt_t<true, true>(a, grid, (unsigned)(lds_bytes + fixed_dt_waves_lds(waves)),
                                         waves, stream);
}

hipError_t fixed_dt_global_window_occupancy(int lds_bytes, int waves, int* blocks_per_cu)
{
    return fixed_dt_occupancy_t<true, true>((size_t)(lds_bytes + fixed_dt_waves_lds(waves)), waves,
                                            blocks_per_cu);
}

}  // namespace mcl
