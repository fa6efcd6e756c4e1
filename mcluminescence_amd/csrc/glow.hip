// glow.hip — glow curves of a sweep on the GPU: the binning and smoothing of the
// reference's plots.py:39-48 (hist_and_smooth) over the data path of plot_TL
// (plots.py:50-66), one curve per run, read straight from the Monte Carlo
// kernels' per-step counts.
//
// One workgroup per run (grid-stride over runs):
//   1. max temperature over the first sim's steps with t > 0 (block max; on the
//      fixed-dt axis the larger of the first and last step's, T being monotone);
//   2. the bin count of np.arange(0, max + w, w) edges, the smoothing width and
//      the curve length, with the too-many-bins / curve-too-long checks;
//   3. per step: the sum of the run's sims' counts (coalesced, four steps per
//      lane when rows allow 16-byte loads) added into a u64 LDS histogram at the
//      bin np.histogram would choose (edges j * w, last bin closed);
//   4. a block scan of the histogram and the 'valid' boxcar as differences of
//      the prefix sums: exact integer window sums, one division each.
// The histogram is sized to the launch's largest run when the host can tell
// (fixed-dt axis, or a host time axis), so small curves leave LDS for more
// workgroups per CU.  Bytes per launch: sims x steps x count width + the first
// sim's time axis + the curves; the histogram never leaves LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mcl.h"
#include "mcl_ctx.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

enum : int32_t { kGlowOk = 0, kGlowTooManyBins = 1, kGlowCurveTooLong = 2, kGlowNoRoom = 3 };

struct GlowRun {
    int64_t first;      // first cluster of the run
    int32_t sims;
    int32_t n_steps;    // steps of the fixed-dt axis (used without a time axis)
    double dt;
    double T_start;
    double T_rate;
};

struct GlowArgs {
    const GlowRun* runs;
    int32_t n_runs;
    int32_t count_bytes;    // 4: u32 counts, 1: u8 lum
    int32_t vec4;           // rows allow four-step vector loads
    int32_t k;              // boxcar width in bins
    int32_t hist_cap;       // histogram bins the launch's LDS holds
    const void* counts;
    int64_t row_stride;
    const double* t_axis;   // [clusters][row_stride] or null (fixed-dt axis)
    double bw;
    double* curve;
    int64_t curve_stride;
    int32_t* curve_len;
    int32_t* status;
};

__device__ __forceinline__ double t_of(const GlowArgs& a, const GlowRun& R, int64_t i)
{
    if (a.t_axis) return a.t_axis[R.first * a.row_stride + i];
    return i < R.n_steps ? (double)(i + 1) * R.dt : 0.0;
}

// Bin of temperature T among edges j * bw, j = 0 .. n_bins: e_j <= T < e_{j+1},
// the last bin closed on the right; -1 outside (np.histogram's rule).
__device__ __forceinline__ int bin_of(double T, double bw, double inv_bw, int n_bins)
{
    if (!(T >= 0.0)) return -1;
    const double q = T * inv_bw;    // a first guess; the edge tests below decide
    int64_t j = q < (double)n_bins ? (int64_t)q : (int64_t)n_bins;
    while (j > 0 && (double)j * bw > T) --j;
    while (j < n_bins && (double)(j + 1) * bw <= T) ++j;
    if (j < n_bins) return (int)j;
    return T == (double)n_bins * bw ? n_bins - 1 : -1;
}

__device__ __forceinline__ double wave_max_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t step_sum(const GlowArgs& a, const GlowRun& R, int64_t i)
{
    uint64_t s = 0;
    if (a.count_bytes == 4) {
        const uint32_t* c = (const uint32_t*)a.counts + R.first * a.row_stride + i;
        for (int q = 0; q < R.sims; ++q) s += c[(int64_t)q * a.row_stride];
    } else {
        const uint8_t* c = (const uint8_t*)a.counts + R.first * a.row_stride + i;
        for (int q = 0; q < R.sims; ++q) s += c[(int64_t)q * a.row_stride];
    }
    return s;
}

// Sums of four consecutive steps i .. i+3 (i a multiple of 4, rows 16-byte aligned).
__device__ __forceinline__ void step_sum4(const GlowArgs& a, const GlowRun& R, int64_t i,
                                          uint64_t (&s)[4])
{
    s[0] = s[1] = s[2] = s[3] = 0;
    if (a.count_bytes == 4) {
        const uint32_t* base = (const uint32_t*)a.counts + R.first * a.row_stride + i;
        int q = 0;
        // sixteen rows in flight per lane: enough bytes outstanding per CU to
        // cover HBM latency at two workgroups per CU
        for (; q + 16 <= R.sims; q += 16) {
            uint4 v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                v[u] = *(const uint4*)(base + (int64_t)(q + u) * a.row_stride);
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                s[0] += v[u].x;
                s[1] += v[u].y;
                s[2] += v[u].z;
                s[3] += v[u].w;
            }
        }
        for (; q + 4 <= R.sims; q += 4) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = *(const uint4*)(base + (int64_t)(q + u) * a.row_stride);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                s[0] += v[u].x;
                s[1] += v[u].y;
                s[2] += v[u].z;
                s[3] += v[u].w;
            }
        }
        for (; q < R.sims; ++q) {
            const uint4 v = *(const uint4*)(base + (int64_t)q * a.row_stride);
            s[0] += v.x;
            s[1] += v.y;
            s[2] += v.z;
            s[3] += v.w;
        }
    } else {
        const uint8_t* base = (const uint8_t*)a.counts + R.first * a.row_stride + i;
        for (int q = 0; q < R.sims; ++q) {
            const uint32_t v = *(const uint32_t*)(base + (int64_t)q * a.row_stride);
            s[0] += v & 0xFFu;
            s[1] += (v >> 8) & 0xFFu;
            s[2] += (v >> 16) & 0xFFu;
            s[3] += v >> 24;
        }
    }
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int o)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, o, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), o, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Largest T over the steps with t > 0 of the fixed-dt axis t_i = (i + 1) dt,
// i < n: T is monotone in i, so it is the first or the last step's.
__host__ __device__ __forceinline__ double fixed_axis_tmax(double T_start, double T_rate, double dt,
                                                           int64_t n)
{
    if (n <= 0) return -INFINITY;
    const double Ta = T_start + T_rate * ((double)1 * dt);
    const double Tb = T_start + T_rate * ((double)n * dt);
    return Tb > Ta ? Tb : Ta;
}

// Histogram bins of the edges 0, w, 2w, ... below max T + w (np.arange).
__host__ __device__ __forceinline__ int64_t glow_bins(double tmax, double bw)
{
    if (!(tmax > -INFINITY)) return 0;
    const double len = ceil((tmax + bw) / bw);
    return len >= 2.0 ? (len < 4.0e18 ? (int64_t)len - 1 : (int64_t)4e18) : 0;
}

__global__ __launch_bounds__(kThreads) void glow_kernel(GlowArgs a)
{
    extern __shared__ uint64_t hist[];              // [a.hist_cap]
    __shared__ uint64_t wsum[kWaves];
    __shared__ double wmax[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int r = blockIdx.x; r < a.n_runs; r += gridDim.x) {
        const GlowRun R = a.runs[r];
        const int64_t n = a.t_axis ? a.row_stride
                                   : (R.n_steps < a.row_stride ? (int64_t)R.n_steps : a.row_stride);
        // 1. max T over steps with t > 0 (-inf when there are none)
        double tmax = -INFINITY;
        if (R.sims > 0 && !a.t_axis) {
            tmax = fixed_axis_tmax(R.T_start, R.T_rate, R.dt, n);
        } else if (R.sims > 0) {
            for (int64_t i = tid; i < n; i += kThreads) {
                const double t = t_of(a, R, i);
                if (t > 0.0) {
                    const double T = R.T_start + R.T_rate * t;
                    tmax = T > tmax ? T : tmax;
                }
            }
            tmax = wave_max_d(tmax);
            if (lane == 0) wmax[wave] = tmax;
            __syncthreads();
            tmax = wmax[0];
            for (int w = 1; w < kWaves; ++w) tmax = wmax[w] > tmax ? wmax[w] : tmax;
        }
        // 2. edges 0, w, 2w, ... below max + w: ceil((max + w) / w) of them
        const int64_t n_bins = glow_bins(tmax, a.bw);
        int64_t L = 0;
        if (n_bins > 0) L = n_bins >= a.k ? n_bins - a.k + 1 : (int64_t)a.k - n_bins + 1;
        int32_t st = kGlowOk;
        if (n_bins > MCL_GLOW_MAX_BINS) st = kGlowTooManyBins;
        else if (n_bins > a.hist_cap) st = kGlowNoRoom;
        else if (L > a.curve_stride) st = kGlowCurveTooLong;
        if (tid == 0) {
            a.status[r] = st;
            a.curve_len[r] = st == kGlowTooManyBins ? 0 : (int32_t)L;
        }
        if (st != kGlowOk || L == 0) {
            __syncthreads();
            continue;
        }
        const int nb = (int)n_bins;
        const double inv_bw = 1.0 / a.bw;
        for (int j = tid; j < nb; j += kThreads) hist[j] = 0;
        __syncthreads();
        // 3. weighted histogram of the step sums
        if (a.vec4) {
            for (int64_t i = (int64_t)tid * 4; i < n; i += (int64_t)kThreads * 4) {
                uint64_t s[4];
                step_sum4(a, R, i, s);
                int jp = -1;
                uint64_t acc = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int j = -1;
                    if (i + u < n) {
                        const double t = t_of(a, R, i + u);
                        if (t > 0.0) j = bin_of(R.T_start + R.T_rate * t, a.bw, inv_bw, nb);
                    }
                    if (j != jp) {
                        if (jp >= 0 && acc) atomicAdd((unsigned long long*)&hist[jp], acc);
                        jp = j;
                        acc = 0;
                    }
                    acc += s[u];
                }
                if (jp >= 0 && acc) atomicAdd((unsigned long long*)&hist[jp], acc);
            }
        } else {
            for (int64_t i = tid; i < n; i += kThreads) {
                const double t = t_of(a, R, i);
                if (!(t > 0.0)) continue;
                const int j = bin_of(R.T_start + R.T_rate * t, a.bw, inv_bw, nb);
                if (j < 0) continue;
                const uint64_t s = step_sum(a, R, i);
                if (s) atomicAdd((unsigned long long*)&hist[j], s);
            }
        }
        __syncthreads();
        // 4. inclusive prefix sums of the histogram: per-thread chunk, then the
        //    chunk totals scanned within each wave and across the four waves
        const int chunk = (nb + kThreads - 1) / kThreads;
        const int j0 = tid * chunk, j1 = j0 + chunk < nb ? j0 + chunk : nb;
        uint64_t run = 0;
        for (int j = j0; j < j1; ++j) {
            run += hist[j];
            hist[j] = run;
        }
        uint64_t incl = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t u = shfl_up_u64(incl, o);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint64_t off = incl - run;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        for (int j = j0; j < j1; ++j) hist[j] += off;
        __syncthreads();
        // boxcar: window sums P[m + k - 1] - P[m - 1]; k > n_bins: every output is the total
        const double denom = (double)R.sims * a.bw * (double)a.k;
        double* out = a.curve + (int64_t)r * a.curve_stride;
        const uint64_t total = hist[nb - 1];
        for (int64_t m = tid; m < L; m += kThreads) {
            uint64_t S;
            if (nb >= a.k) {
                const uint64_t hi = hist[m + a.k - 1];
                const uint64_t lo = m > 0 ? hist[m - 1] : 0;
                S = hi - lo;
            } else {
                S = total;
            }
            out[m] = (double)S / denom;
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" int mcl_glow_curve(void* ctx, const mcl_run_params* runs, int32_t n_runs,
                              const void* counts, int32_t count_bytes, int64_t row_stride,
                              const double* t_axis, double bin_width, double win, uint32_t flags,
                              double* curve, int64_t curve_stride, int32_t* curve_len,
                              mcl_stats* stats)
{
    mcl::Ctx* c = (mcl::Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!runs || !counts || !curve || !curve_len)
        return mcl::fail(c, MCL_EINVAL, "runs, counts, curve and curve_len must not be null");
    if (n_runs <= 0) return mcl::fail(c, MCL_EINVAL, "n_runs must be > 0");
    if (count_bytes != 4 && count_bytes != 1)
        return mcl::fail(c, MCL_EINVAL, "count_bytes must be 4 (u32 counts) or 1 (u8 lum)");
    if (row_stride <= 0) return mcl::fail(c, MCL_EINVAL, "row_stride must be > 0");
    if (curve_stride <= 0) return mcl::fail(c, MCL_EINVAL, "curve_stride must be > 0");
    if (!(bin_width > 0.0) || !std::isfinite(bin_width))
        return mcl::fail(c, MCL_EINVAL, "bin_width must be finite and > 0 (got %g)", bin_width);
    if (!(win >= 0.0) || !std::isfinite(win))
        return mcl::fail(c, MCL_EINVAL, "win must be finite and >= 0 (got %g)", win);
    // k = max(1, int(win / bin_width)) (plots.py:46-47)
    const double kq = std::trunc(win / bin_width);
    if (kq > 1.0e9) return mcl::fail(c, MCL_EINVAL, "smoothing window of %g bins is too wide", kq);
    const int32_t k = kq < 1.0 ? 1 : (int32_t)kq;
    std::vector<GlowRun> gr((size_t)n_runs);
    int64_t total = 0;
    for (int r = 0; r < n_runs; ++r) {
        const mcl_run_params& p = runs[r];
        if (p.sims < 0) return mcl::fail(c, MCL_EINVAL, "run %d: sims must be >= 0", r);
        if (!std::isfinite(p.T_start) || !std::isfinite(p.T_rate))
            return mcl::fail(c, MCL_EINVAL, "run %d: T_start and T_rate must be finite", r);
        if (!t_axis && (p.n_steps < 0 || !(p.dt > 0.0) || !std::isfinite(p.dt)))
            return mcl::fail(c, MCL_EINVAL, "run %d: the fixed-dt axis needs dt > 0 and "
                             "n_steps >= 0", r);
        gr[(size_t)r] = GlowRun{total, p.sims, p.n_steps, p.dt, p.T_start, p.T_rate};
        total += p.sims;
    }
    if (total <= 0) return mcl::fail(c, MCL_EINVAL, "no clusters (sum of sims is 0)");
    const bool dev = (flags & MCL_F_DEVICE_PTRS) != 0;
    // histogram bins of the largest run, when the time axis is known here (the
    // kernel's own arithmetic, so the same count); a device time axis gets the most
    int64_t cap = 1;
    if (dev && t_axis) {
        cap = MCL_GLOW_MAX_BINS;
    } else {
        for (int r = 0; r < n_runs; ++r) {
            const GlowRun& g = gr[(size_t)r];
            if (g.sims <= 0) continue;
            double tmax = -INFINITY;
            if (!t_axis) {
                tmax = fixed_axis_tmax(g.T_start, g.T_rate, g.dt,
                                       g.n_steps < row_stride ? (int64_t)g.n_steps : row_stride);
            } else {
                const double* row = t_axis + g.first * row_stride;
                for (int64_t i = 0; i < row_stride; ++i) {
                    if (!(row[i] > 0.0)) continue;
                    const double T = g.T_start + g.T_rate * row[i];
                    tmax = T > tmax ? T : tmax;
                }
            }
            const int64_t nbins = glow_bins(tmax, bin_width);
            cap = std::max(cap, std::min(nbins, (int64_t)MCL_GLOW_MAX_BINS));
        }
    }
    const size_t lds_bytes = sizeof(uint64_t) * (size_t)cap;
    if (hipSetDevice(c->device) != hipSuccess)
        return mcl::fail(c, MCL_EDEVICE, "hipSetDevice failed");
    const size_t cnt_bytes = (size_t)total * (size_t)row_stride * (size_t)count_bytes;
    const size_t t_bytes = (size_t)total * (size_t)row_stride * sizeof(double);
    const size_t out_bytes = (size_t)n_runs * (size_t)curve_stride * sizeof(double);
    std::vector<void*> tmp;
    auto dmalloc = [&](void** ptr, size_t bytes) {
        const hipError_t e = hipMalloc(ptr, bytes);
        if (e == hipSuccess) tmp.push_back(*ptr);
        return e;
    };
    auto release = [&]() {
        for (void* q : tmp) (void)hipFree(q);
    };
    GlowRun* d_runs = nullptr;
    int32_t* d_status = nullptr;
    const void* d_counts = dev ? counts : nullptr;
    const double* d_t = dev ? t_axis : nullptr;
    double* d_curve = dev ? curve : nullptr;
    int32_t* d_len = dev ? curve_len : nullptr;
    hipError_t e = dmalloc((void**)&d_runs, sizeof(GlowRun) * gr.size());
    if (e == hipSuccess) e = dmalloc((void**)&d_status, sizeof(int32_t) * (size_t)n_runs);
    if (e == hipSuccess && !dev) e = dmalloc((void**)&d_counts, cnt_bytes);
    if (e == hipSuccess && !dev && t_axis) e = dmalloc((void**)&d_t, t_bytes);
    if (e == hipSuccess && !dev) e = dmalloc((void**)&d_curve, out_bytes);
    if (e == hipSuccess && !dev) e = dmalloc((void**)&d_len, sizeof(int32_t) * (size_t)n_runs);
    if (e != hipSuccess) {
        release();
        return mcl::fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE,
                         "allocation: %s", hipGetErrorString(e));
    }
    e = hipMemcpyAsync(d_runs, gr.data(), sizeof(GlowRun) * gr.size(), hipMemcpyHostToDevice,
                       c->stream);
    if (e == hipSuccess && !dev)
        e = hipMemcpyAsync((void*)d_counts, counts, cnt_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !dev && t_axis)
        e = hipMemcpyAsync((void*)d_t, t_axis, t_bytes, hipMemcpyHostToDevice, c->stream);
    GlowArgs a{};
    a.runs = d_runs;
    a.n_runs = n_runs;
    a.count_bytes = count_bytes;
    a.vec4 = (row_stride % 4 == 0) && ((uintptr_t)d_counts % 16 == 0);
    a.k = k;
    a.hist_cap = (int32_t)cap;
    a.counts = d_counts;
    a.row_stride = row_stride;
    a.t_axis = d_t;
    a.bw = bin_width;
    a.curve = d_curve;
    a.curve_stride = curve_stride;
    a.curve_len = d_len;
    a.status = d_status;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) e = hipEventRecord(ev0, c->stream);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)glow_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e == hipSuccess) {
        const int grid = n_runs < 65536 ? n_runs : 65536;
        hipLaunchKernelGGL(glow_kernel, dim3(grid), dim3(kThreads), lds_bytes, c->stream, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev1, c->stream);
    std::vector<int32_t> hst((size_t)n_runs), hlen((size_t)n_runs);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hst.data(), d_status, sizeof(int32_t) * (size_t)n_runs,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hlen.data(), d_len, sizeof(int32_t) * (size_t)n_runs,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !dev)
        e = hipMemcpyAsync(curve, d_curve, out_bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ev0, ev1);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    release();
    if (e != hipSuccess)
        return mcl::fail(c, MCL_EDEVICE, "glow-curve kernel or copy failed: %s",
                         hipGetErrorString(e));
    if (!dev) std::memcpy(curve_len, hlen.data(), sizeof(int32_t) * (size_t)n_runs);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->kernel_ms = ms;
        stats->clusters = (uint64_t)total;
    }
    for (int r = 0; r < n_runs; ++r) {
        if (hst[(size_t)r] == kGlowTooManyBins)
            return mcl::fail(c, MCL_ERANGE, "run %d: the histogram has more than %d bins of "
                             "width %g", r, MCL_GLOW_MAX_BINS, bin_width);
        if (hst[(size_t)r] == kGlowNoRoom)
            return mcl::fail(c, MCL_EDEVICE, "run %d: histogram larger than the launch's LDS "
                             "(%lld bins)", r, (long long)cap);
        if (hst[(size_t)r] == kGlowCurveTooLong)
            return mcl::fail(c, MCL_EOVERFLOW, "run %d: the curve has %d points; curve_stride "
                             "is %lld", r, hlen[(size_t)r], (long long)curve_stride);
    }
    return MCL_OK;
}
