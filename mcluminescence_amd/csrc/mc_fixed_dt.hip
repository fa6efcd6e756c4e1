// mc_fixed_dt.hip — the fixed-dt kernels whose clusters live wholly in LDS (the
// cluster body, its layout and the kernel templates are in mc_fixed_dt.h).
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
    hipError_t e = hipSuccess;
    for (int w : {1, 2, 4, 8}) {
        for (bool fw : {false, true}) {
            if (e == hipSuccess)
                e = hipFuncSetAttribute(fixed_dt_fn<false>(w, fw),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        std::min(lds_bytes + fixed_dt_waves_lds(w), max_lds));
        }
    }
    return e;
}

hipError_t launch_fixed_dt(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                           hipStream_t stream)
{
    return launch_fixed_dt_t<false>(a, grid, (unsigned)(lds_bytes + fixed_dt_waves_lds(waves)),
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
    return fixed_dt_occupancy_t<false>((size_t)(lds_bytes + fixed_dt_waves_lds(waves)), waves,
                                       blocks_per_cu);
}

}  // namespace mcl
