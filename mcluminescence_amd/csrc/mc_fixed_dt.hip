// mc_fixed_dt.hip — the fixed-dt kernels whose clusters live wholly in LDS (the
// cluster body, its layout and the kernel templates are in mc_fixed_dt.h), the
// builds without the quiet window; mc_fixed_dt_window.hip holds those with it.
#include "mc_fixed_dt.h"

#include <algorithm>

namespace mcl {

// Zero the tickets of the persistent class kernels (a kernel node, so the
// graph holds no memset node).
__global__ void reset_tickets_kernel(uint32_t* tickets, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) tickets[i] = 0u;
}

hipError_t prepare_fixed_dt(int lds_bytes, int max_lds)
{
    hipError_t e = prepare_fixed_dt_t<false, false>(lds_bytes, max_lds);
    if (e == hipSuccess) e = prepare_fixed_dt_window(lds_bytes, max_lds);
    return e;
}

hipError_t launch_fixed_dt(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                           hipStream_t stream)
{
    if (a.quiet_window != 0) return launch_fixed_dt_window(a, grid, lds_bytes, waves, stream);
    return launch_fixed_dt_t<false, false>(a, grid, (unsigned)(lds_bytes + fixed_dt_waves_lds(waves)),
                                           waves, stream);
}

hipError_t launch_reset_tickets(uint32_t* tickets, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_tickets_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0u, stream,
                       tickets, n);
    return hipGetLastError();
}

hipError_t fixed_dt_occupancy(int lds_bytes, int waves, int* blocks_per_cu)
{
    int b0 = 0, b1 = 0;
    hipError_t e = fixed_dt_occupancy_t<false, false>((size_t)(lds_bytes + fixed_dt_waves_lds(waves)),
                                                      waves, &b0);
    if (e == hipSuccess) e = fixed_dt_window_occupancy(lds_bytes, waves, &b1);
    *blocks_per_cu = b0 < b1 ? b0 : b1;
    return e;
}

}  // namespace mcl
