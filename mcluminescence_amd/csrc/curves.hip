// curves.hip — per-set curves from the Monte Carlo's per-step counts: the sum over
// each parameter set's sims and over the steps of each bin, with bin edges given
// in steps (the surrogate's log-time isothermal decay curves: single steps
// first, geometrically wider bins later).  The reference has no counterpart; its
// nearest relative is the histogram of plots.py:39-47, which bins by
// temperature (glow.hip).
//
// One workgroup per parameter set (grid-stride over sets):
//   1. per step, the sum of the set's sims' counts into an LDS row (sim rows read
//      coalesced, four steps per lane with 16-byte loads when rows allow);
//   2. per bin, one lane sums its steps of that row (edges are in the kernel
//      arguments, at most MCL_BIN_MAX + 1 of them).
// HBM-bound: sims x n_steps x 4 B read per set, 8 B per bin written.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mcl.h"
#include "mcl_ctx.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxSteps = 16384;          // LDS row of step totals: 64 KiB

struct BinEdges {
    int32_t e[MCL_BIN_MAX + 1];
};

struct BinArgs {
    const uint32_t* counts;   // [n_sets * sims][row_stride]
    int64_t n_sets;
    int64_t row_stride;
    int32_t sims;
    int32_t n_steps;          // steps read per row (the last edge)
    int32_t n_bins;
    int32_t vec4;             // row_stride and the base allow 16-byte loads
    uint64_t* out;            // [n_sets][n_bins]
};

__global__ __launch_bounds__(kThreads) void bin_counts_kernel(BinArgs a, BinEdges ed)
{
    __shared__ uint32_t tot[kMaxSteps];
    const int tid = threadIdx.x;
    for (int64_t set = blockIdx.x; set < a.n_sets; set += gridDim.x) {
        const uint32_t* rows = a.counts + set * (int64_t)a.sims * a.row_stride;
        if (a.vec4) {
            const int nq = (a.n_steps + 3) / 4;
            for (int q = tid; q < nq; q += kThreads) {
                uint4 acc = make_uint4(0u, 0u, 0u, 0u);
                for (int k = 0; k < a.sims; ++k) {
                    const uint4 v = *(const uint4*)(rows + k * a.row_stride + 4 * (int64_t)q);
                    acc.x += v.x;
                    acc.y += v.y;
                    acc.z += v.z;
                    acc.w += v.w;
                }
                // steps past n_steps (up to the next multiple of 4) are in the row
                // (row_stride is a multiple of 4 >= that) and never read by a bin
                tot[4 * q] = acc.x;
                tot[4 * q + 1] = acc.y;
                tot[4 * q + 2] = acc.z;
                tot[4 * q + 3] = acc.w;
            }
        } else {
            for (int s = tid; s < a.n_steps; s += kThreads) {
                uint32_t acc = 0u;
                for (int k = 0; k < a.sims; ++k) acc += rows[k * a.row_stride + s];
                tot[s] = acc;
            }
        }
        __syncthreads();
        for (int j = tid; j < a.n_bins; j += kThreads) {
            uint64_t sum = 0u;
            for (int s = ed.e[j]; s < ed.e[j + 1]; ++s) sum += tot[s];
            a.out[set * a.n_bins + j] = sum;
        }
        __syncthreads();            // the row is rewritten by the next set
    }
}

}  // namespace

extern "C" int mcl_bin_counts(void* ctx, const uint32_t* counts_dev, int64_t n_sets, int32_t sims,
                              int64_t row_stride, const int32_t* edges, int32_t n_bins,
                              uint64_t* out_dev)
{
    mcl::Ctx* c = (mcl::Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!counts_dev || !out_dev || !edges) return mcl::fail(c, MCL_EINVAL, "null pointer");
    if (n_sets <= 0 || sims <= 0) return mcl::fail(c, MCL_EINVAL, "n_sets and sims must be > 0");
    if (n_bins <= 0 || n_bins > MCL_BIN_MAX)
        return mcl::fail(c, MCL_ERANGE, "n_bins=%d must be in [1, %d]", n_bins, MCL_BIN_MAX);
    BinEdges ed;
    if (edges[0] < 0) return mcl::fail(c, MCL_EINVAL, "edges[0]=%d is negative", edges[0]);
    for (int j = 0; j <= n_bins; ++j) {
        ed.e[j] = edges[j];
        if (j > 0 && edges[j] <= edges[j - 1])
            return mcl::fail(c, MCL_EINVAL, "edges must increase (edges[%d]=%d, edges[%d]=%d)",
                             j - 1, edges[j - 1], j, edges[j]);
    }
    const int32_t n_steps = edges[n_bins];
    if (n_steps > kMaxSteps || (int64_t)n_steps > row_stride)
        return mcl::fail(c, MCL_ERANGE, "last edge %d exceeds the row (%lld) or %d steps", n_steps,
                         (long long)row_stride, kMaxSteps);
    BinArgs a;
    a.counts = counts_dev;
    a.n_sets = n_sets;
    a.row_stride = row_stride;
    a.sims = sims;
    a.n_steps = n_steps;
    a.n_bins = n_bins;
    a.vec4 = (row_stride % 4 == 0) && ((uintptr_t)counts_dev % 16 == 0) &&
             ((int64_t)((n_steps + 3) / 4) * 4 <= row_stride);
    a.out = out_dev;
    if (hipSetDevice(c->device) != hipSuccess) return mcl::fail(c, MCL_EDEVICE, "hipSetDevice failed");
    const int64_t grid = n_sets < 65536 ? n_sets : 65536;
    hipLaunchKernelGGL(bin_counts_kernel, dim3((unsigned)grid), dim3(kThreads), 0, c->stream, a, ed);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mcl::fail(c, MCL_EDEVICE, "bin_counts launch: %s", hipGetErrorString(e));
    return MCL_OK;
}
