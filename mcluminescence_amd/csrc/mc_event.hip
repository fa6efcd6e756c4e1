// mc_event.hip — event-mode trap/charge-kinetics Monte Carlo, one wave64 per
// cluster, one event per step: the reference's loop (simulate.py:46-92 driving
// TLTrapSim._update_lifetimes / _filling_time, tl_trap_lab.py:48-60, and
// Box.remove_pair, engine.py:154-175), for the exact step / lum / electron_ratio
// semantics of its CSV.  Geometry and nearest holes are those of the fixed-dt
// kernel (mc_common.h).
//
// Step i, with t the time so far (float64):
//   T = (T_start + T_rate t) + 273.15 (simulate.py:53-54); A = s e^{-E_cb/kT} and
//   B_c = b e^{-E_loc_c/kT} in float64 with the library's exp; per live electron
//   (index order) k_e = fma(B_ch, w_e, A) in float64, the channel drawn as in the
//   fixed-dt step (Philox (i >> 2, e, run, sim|CHAN), word i & 3).
//   K = sum of k_e: each row of 64 consecutive live electrons is summed as the
//   last lane of an inclusive scan (row_scan_d: steps of 1, 2, 4, 8 within 16-lane
//   rows, then the totals of the rows before; row_total_d forms the same value),
//   and the rows are added in order.
//   One Philox block (i, 0xFFFFFFFF, run, sim|FIRE) gives three words:
//   E_t = -ln((u0 + 1/2) 2^-32), U_w = u1 2^-32, E_f = -ln((u2 + 1/2) 2^-32).
//   dt_fill = E_f (1 / lambda), lambda = 1e-20 when D = 0 or n_e = N_e
//   (tl_trap_lab.py:53-60); dt_rec = E_t / K: the earliest of independent
//   exponential waits with rates k_e (tl_trap_lab.py:48-51) is exponential with
//   rate K; with no electron dt_rec = dt_fill, and +inf when every rate is 0
//   (simulate.py:59).  dt = min(dt_fill, dt_rec, dt_cap), dt_cap = max_dt / T_rate,
//   or 1e20 without a ramp (simulate.py:45,60).
//   t += dt.  dt == dt_fill: a filling event (only the terminal one of an emptied
//   isothermal run exists without dose); else dt == dt_rec: the recombining
//   electron is the first whose running sum exceeds U_w K, i.e. electron e with
//   probability k_e / K, the law of the argmin of the waits.  It and its nearest
//   hole are removed, the live electrons are compacted (index order kept), and
//   the electrons that had that hole are searched again.
//   Row i: lum = 1 for a recombination, else 0; t; n_e / N_e.  The loop runs while
//   t <= duration and stops after a step that reaches duration (simulate.py:51,91).
// The oracle (oracle/mcl_oracle.c, orc_run_event) replays the same arithmetic in
// the same order, so every row is bit-exact.
//
// LDS: hole [h_cap] and cstart [g_cap] as in the fixed-dt kernel, then est, ew,
// the row sums and rq [e_cap each].  Clusters beyond LDS take the fixed-dt
// kernel's global-memory form (GM): hole and cstart in the workgroup's scratch
// slice, the electron arrays in LDS while they fit (e_lds) or after cstart in the
// slice, persistent workgroups over a ticket; the wave's fences widen to
// workgroup scope so its global stores complete before its lanes read them.
#include "mc_common.h"

namespace mcl {


__device__ __forceinline__ double exp1_of(uint32_t u)
{
    return -log_own(((double)u + 0.5) * 0x1p-32);
}

// An upper bound of boltzmann(pre, pre_ok, E, kT) for a step that may skip K:
// 0 where boltzmann is 0 by its own test; otherwise e^{-E/kT} in float32 (the
// quotient from a reciprocal and the hardware exp2, exponent clamped at -80 so the
// result stays a normal float), whose errors, below 1e-4 relative, the factor 1 +
// 2^-10 covers; +inf (no skip) when pre is not finite and positive.
__device__ __forceinline__ double boltzmann_ub(double pre, bool pre_ok, double E, double kT)
{
    if (!pre_ok) return __builtin_inf();
    if (kT * 720.0 < E) return 0.0;
    float x = -(float)E * __builtin_amdgcn_rcpf((float)kT);
    x = fmaxf(x, -80.0f);
    const float e = __builtin_amdgcn_exp2f(x * 1.44269504f);
    return pre * (double)e * (1.0 + 0x1p-10);
}

// pre e^{-E/kT} as the step computes it (the library's exp of the float64
// quotient), with the quotient and exp skipped where kT 720 < E: there E / kT >
// 708, where exp_own returns 0, so the factor is 0 either way (pre_ok: pre finite
// and > 0).  The tunnelling physics' band (E_cb = 1000 eV) is such a term at
// every temperature.
__device__ __forceinline__ double boltzmann(double pre, bool pre_ok, double E, double kT)
{
    if (pre_ok && kT * 720.0 < E) return 0.0;
    return pre * exp_own(-(E / kT));
}

// Ordering of one wave's writes before its lanes read them: a wavefront-scope
// fence for LDS, workgroup scope (the wave's global stores complete) for GM.
template <bool GM>
__device__ __forceinline__ void event_fence()
{
    if constexpr (GM) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        wave_lds_fence();
    }
}

template <bool GM>
__device__ __forceinline__ void run_cluster_event(const EventArgs& a, const int c,
                                                  unsigned char* lds_raw)
{
    const int lane = (int)threadIdx.x;
    const int run = uniform(a.c_run[c]);
    const int sim = uniform(a.c_sim[c]);
    const RunDesc R = a.runs[run];
    const int n_e0 = uniform(R.n_e0);
    const int h_cap = a.h_cap, e_cap = a.e_cap, g_cap = a.g_cap;

    unsigned char* hbase = lds_raw;     // hole, cstart
    unsigned char* ebase = lds_raw + 16 * h_cap + 4 * g_cap;   // est, ew, row sums, rq
    if constexpr (GM) {
        unsigned char* const slice = a.gscratch + (size_t)blockIdx.x * (size_t)a.gstride;
        hbase = slice;
        ebase = a.e_lds ? lds_raw : slice + 16 * (size_t)h_cap + 4 * (size_t)g_cap;
    }
    float4* hole = (float4*)hbase;
    float* hole_f = (float*)hbase;
    uint32_t* cstart = (uint32_t*)(hole + h_cap);
    uint32_t* est = (uint32_t*)ebase;
    float* ew = (float*)(est + e_cap);
    double* rtot = (double*)(ew + e_cap);            // row sums (the ethr words)
    uint32_t* rq = (uint32_t*)(ew + e_cap) + e_cap;

    const uint32_t k0 = a.key0, k1 = a.key1;
    const uint32_t prun = (uint32_t)R.run_id;
    const uint32_t psim = (uint32_t)sim << 4;
    const ElectronGen egen{prun, psim | MCL_TAG_GEOM_E, k0, k1, R.side_nm};

    uint64_t dist_evals = 0;
    const Grid grid = setup_cluster(R, egen, hole, cstart, est, ew, prun, psim, k0, k1, lane,
                                    dist_evals);
    const float alpha_f = (float)R.alpha_nm;

    const bool two_ch = R.two_channel != 0;
    const bool always2 = R.always_chan2 != 0;
    const uint64_t ret_thr = R.ret_thr;
    const uint32_t c3f = psim | MCL_TAG_FIRE, c3c = psim | MCL_TAG_CHAN;
    const bool b1_is_a = R.b == R.s && R.E_loc_1 == R.E_cb;
    const bool b2_is_b1 = R.E_loc_2 == R.E_loc_1;
    const double dt_cap = R.T_rate != 0.0 ? R.dt / R.T_rate : 1e20;
    const int N_int = (int)R.N_e;
    const int max_steps = a.max_steps;
    uint8_t* lum = a.lum + (size_t)c * (size_t)max_steps;
    double* tax = a.t_axis ? a.t_axis + (size_t)c * (size_t)max_steps : nullptr;
    double* erat = a.e_ratio ? a.e_ratio + (size_t)c * (size_t)max_steps : nullptr;

    int n_live = n_e0;        // live electrons occupy slots 0 .. n_live - 1
    int filled = 0;
    double t = 0.0;
    int i = 0;
    int32_t status = 0;
    uint64_t lsteps = 0, recomb = 0;
    const bool s_ok = R.s > 0.0 && R.s < __builtin_inf();
    const bool b_ok = R.b > 0.0 && R.b < __builtin_inf();
    // Without dose, lambda is 1e-20 and dt_fill >= 1e20 E_f >= 1.1e10 s (E_f >=
    // -ln(1 - 2^-33)): with dt_cap below that it is never the minimum, so the
    // step does not draw it (dt_fill = +inf gives the same dt and the same event).
    const bool fill_never = R.D == 0.0 && dt_cap < 1e10;
    int erat_n = -1;          // n_e of the last e_ratio value computed
    double erat_v = 0.0;
    // The step draws depend on the step index only: lane j holds those of step
    // sb_base + j (U_w, E_t and E_f, each computed as the step would), 64 steps
    // in one parallel evaluation; a step reads its own with scalar reads.
    int sb_base = -kWave;
    uint32_t l_uw = 0u;
    double l_et = 0.0, l_ef = 0.0;
    // The rows go out 64 at a time: lane j keeps row out_base + j until the
    // block is full or the loop ends, then every lane stores its own (coalesced).
    int out_base = 0;
    uint32_t o_lum = 0u;
    double o_t = 0.0, o_er = 0.0;
    // An upper bound of the sum of w over the live electrons that have a hole,
    // valid until the next recombination changes the electrons or their w.
    bool w_valid = false;
    double w_ub = 0.0;
    auto flush_rows = [&](int upto) {
        const int row = out_base + lane;
        if (row < upto) {
            lum[row] = (uint8_t)o_lum;
            if (tax) tax[row] = o_t;
            if (erat) erat[row] = o_er;
        }
        out_base = upto;
    };

    while (t <= R.duration) {
        if (i >= max_steps) {
            status = -1 - i;
            break;
        }
        lsteps += (uint64_t)n_live;
        const double T = (R.T_start_C + R.T_rate * t) + 273.15;
        const double kT = R.k_b * T;
        double A = 0.0, B1 = 0.0, B2 = 0.0;   // the exact factors, for a step that forms K
        const uint32_t blk = (uint32_t)i >> 2;
        const int word = i & 3;
        // k_e of slot p, 0 for slots past the live ones and electrons with no hole;
        // branch-free: every slot of a row lies within e_cap, so the reads are in
        // bounds, and the selects drop what the slots past n_live hold
        auto rate_of = [&](int p) -> double {
            const uint32_t st = est[p];
            const float w = ew[p];
            bool ch = always2;
            if (two_ch)
                ch = (uint64_t)pick_word(philox10(blk, st & 0xFFFFu, prun, c3c, k0, k1), word) <
                     ret_thr;
            const double k = __builtin_fma(ch ? B2 : B1, (double)w, A);
            return p < n_live && (st >> 16) != kNoHole ? k : 0.0;
        };
        if (i - sb_base >= kWave) {
            const u32x4 o = philox10((uint32_t)(i + lane), 0xFFFFFFFFu, prun, c3f, k0, k1);
            l_uw = o.y;
            l_et = exp1_of(o.x);
            l_ef = fill_never ? 0.0 : exp1_of(o.z);
            sb_base = i;
        }
        const int js = i - sb_base;
        const uint32_t u_w = lane_value(l_uw, js);
        const double e_t = lane_value_d(l_et, js);
        const double lam = (R.D == 0.0 || n_live + filled == N_int)
                               ? 1e-20
                               : (R.D / R.D0) * (double)(N_int - (n_live + filled));
        const double dt_fill =
            fill_never ? __builtin_inf() : (1.0 / lam) * lane_value_d(l_ef, js);
        // A step whose cap is certainly the minimum needs no K: every k_e is at most
        // max(B1, B2) w_e + A (rounded up by at most an ulp), so K, summed in
        // float64 over at most 65535 terms, is below K_ub = (B_max W + n A)(1 +
        // 2^-30), W an upper bound of the sum of w; then E_t > dt_cap K_ub (1 +
        // 2^-30) gives E_t / K > dt_cap in float64 as well, and with dt_fill >
        // dt_cap the step takes dt = dt_cap and no event, as the full step would.
        bool quiet = false;
        if (n_live > 0 && dt_fill > dt_cap) {
            if (!w_valid) {
                // (any summation order serves a bound: two rows at a time, then the
                // in-row DPP steps and scalar reads of row_total_d)
                double ws = 0.0;
                int p = lane;
                for (; p + kWave < n_live; p += 2 * kWave) {
                    const uint32_t sa = est[p], sb2 = est[p + kWave];
                    const float wa = ew[p], wb = ew[p + kWave];
                    ws += ((sa >> 16) != kNoHole ? (double)wa : 0.0) +
                          ((sb2 >> 16) != kNoHole ? (double)wb : 0.0);
                }
                if (p < n_live) {
                    const uint32_t sa = est[p];
                    const float wa = ew[p];
                    ws += (sa >> 16) != kNoHole ? (double)wa : 0.0;
                }
                w_ub = row_total_d(ws) * (1.0 + 0x1p-30);
                w_valid = true;
            }
            // bounds of the factors (not the factors: a step that skips K needs no more)
            const double a_ub = boltzmann_ub(R.s, s_ok, R.E_cb, kT);
            const double b1_ub = b1_is_a ? a_ub : boltzmann_ub(R.b, b_ok, R.E_loc_1, kT);
            const double b2_ub = b2_is_b1 ? b1_ub : boltzmann_ub(R.b, b_ok, R.E_loc_2, kT);
            const double bm = two_ch ? fmax(b1_ub, b2_ub) : (always2 ? b2_ub : b1_ub);
            const double k_ub = (bm * w_ub + (double)n_live * a_ub) * (1.0 + 0x1p-30);
            quiet = e_t > dt_cap * k_ub * (1.0 + 0x1p-30);
        }
        const int rows = (n_live + kWave - 1) / kWave;
        double K = 0.0;
        if (!quiet) {
            A = boltzmann(R.s, s_ok, R.E_cb, kT);
            B1 = b1_is_a ? A : boltzmann(R.b, b_ok, R.E_loc_1, kT);
            B2 = b2_is_b1 ? B1 : boltzmann(R.b, b_ok, R.E_loc_2, kT);
            // row totals (the scan's lane-63 values), two rows at a time (independent
            // chains), added in order
            int r = 0;
            for (; r + 1 < rows; r += 2) {
                const double ra = row_total_d(rate_of(r * kWave + lane));
                const double rb = row_total_d(rate_of((r + 1) * kWave + lane));
                if (lane == 0) {
                    rtot[r] = ra;
                    rtot[r + 1] = rb;
                }
                K = K + ra;
                K = K + rb;
            }
            if (r < rows) {
                const double rt = row_total_d(rate_of(r * kWave + lane));
                if (lane == 0) rtot[r] = rt;
                K = K + rt;
            }
            event_fence<GM>();
        }
        double dt_rec = dt_fill;
        if (n_live > 0) dt_rec = quiet ? __builtin_inf() : (K > 0.0 ? e_t / K : __builtin_inf());
        const double dt = fmin(fmin(dt_fill, dt_rec), dt_cap);
        t += dt;
        uint8_t ev = 0;
        if (dt == dt_fill) {
            filled += 1;
            if (!(t >= R.duration && R.duration != 0.0)) {
                status = kEventFillUnsupported;   // the new pair would take part
                break;
            }
        } else if (dt == dt_rec) {
            // the recombining electron: first running sum above U_w K
            const double target = (double)u_w * 0x1p-32 * K;
            int pw = -1;
            double carry = 0.0;
            for (int r = 0; r < rows; ++r) {
                const double c2 = carry + rtot[r];
                if (c2 > target) {
                    const double v = row_scan_d(rate_of(r * kWave + lane), lane);
                    const uint64_t m = ballot(carry + v > target);
                    pw = r * kWave + (int)__builtin_ctzll(m);
                    break;
                }
                carry = c2;
            }
            if (pw < 0) {
                // U_w K rounded up to K: the last electron with a rate
                for (int r = rows - 1; r >= 0 && pw < 0; --r) {
                    const uint64_t m = ballot(rate_of(r * kWave + lane) > 0.0);
                    if (m != 0) pw = r * kWave + 63 - (int)__builtin_clzll(m);
                }
            }
            pw = uniform(pw);
            const uint32_t sw = est[pw];
            const uint32_t hw = sw >> 16;
            if (lane == 0) hole_f[4 * hw] = kDeadHoleX;
            // One pass over the rows: slots pw + 1 .. n_live - 1 move down by one (the
            // index order kept; a row reads its sources before it writes, and the next
            // row's sources lie beyond what it wrote), and the electrons whose nearest
            // hole was hw are queued, at their new slots, for a new search
            // (engine.py:168-175).
            const int n_new = n_live - 1;
            int nq = 0;
            for (int r = 0; r * kWave < n_new; ++r) {
                const int p = r * kWave + lane;
                const bool in = p < n_new, moves = in && p >= pw;
                const uint32_t st = in ? est[moves ? p + 1 : p] : kDeadSlot;
                const float w = moves ? ew[p + 1] : 0.0f;
                if (moves) {
                    est[p] = st;
                    ew[p] = w;
                }
                const bool s = (st >> 16) == hw;
                const uint64_t m = ballot(s);
                if (s) rq[nq + lanes_below(m)] = ((uint32_t)p << 16) | (st & 0xFFFFu);
                nq += __popcll(m);
            }
            n_live = n_new;
            recomb += 1;
            w_valid = false;
            event_fence<GM>();
            if (nq > 0) {
                search_queue_nn(grid, egen, est, ew, rq, nq, a.wave_search_max, alpha_f, lane,
                                dist_evals);
                event_fence<GM>();
            }
            ev = 1;
        }
        if (erat && n_live + filled != erat_n) {
            erat_n = n_live + filled;
            erat_v = (double)erat_n / R.N_e;
        }
        if (lane == i - out_base) {
            o_lum = ev;
            o_t = t;
            o_er = erat_v;
        }
        ++i;
        if (i - out_base == kWave) flush_rows(i);
        if (R.duration != 0.0 && t >= R.duration) break;
    }
    const int used = i;
    flush_rows(used);
    for (int r = used + lane; r < max_steps; r += kWave) {
        lum[r] = 0;
        if (tax) tax[r] = 0.0;
        if (erat) erat[r] = 0.0;
    }
    uint64_t evals_total = dist_evals;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)evals_total, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(evals_total >> 32), off);
        evals_total += ((uint64_t)hi << 32) | lo;
    }
    if (lane == 0) {
        a.steps[c] = status != 0 ? status : used;
        if (a.work != nullptr) {
            a.work[4 * (size_t)c] = lsteps;
            a.work[4 * (size_t)c + 1] = evals_total;
            a.work[4 * (size_t)c + 2] = recomb;
            a.work[4 * (size_t)c + 3] = (uint64_t)used;
        }
    }
}

// One workgroup per cluster, or (queue set: the global-memory form) persistent
// workgroups taking clusters 0 .. n_clusters-1 by ticket until it passes the end.
template <bool GM>
__global__ __launch_bounds__(64) void event_kernel(EventArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (a.queue == nullptr) {
        if ((int)blockIdx.x < a.n_clusters) run_cluster_event<GM>(a, (int)blockIdx.x, lds_raw);
        return;
    }
    for (;;) {
        int i = 0;
        if (threadIdx.x == 0) i = (int)atomicAdd(a.queue, 1u);
        i = __shfl(i, 0);
        if (i >= a.n_clusters) break;
        run_cluster_event<GM>(a, i, lds_raw);
        __syncthreads();
    }
}

hipError_t prepare_event(int lds_bytes)
{
    hipError_t e = hipFuncSetAttribute((const void*)event_kernel<false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)event_kernel<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    return e;
}

hipError_t launch_event(const EventArgs& a, int lds_bytes, hipStream_t stream)
{
    hipLaunchKernelGGL(event_kernel<false>, dim3((unsigned)a.n_clusters), dim3(kWave),
                       (unsigned)lds_bytes, stream, a);
    return hipGetLastError();
}

hipError_t launch_event_global(const EventArgs& a, int grid, int lds_bytes, hipStream_t stream)
{
    if (a.queue == nullptr || a.gscratch == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(event_kernel<true>, dim3((unsigned)grid), dim3(kWave), (unsigned)lds_bytes,
                       stream, a);
    return hipGetLastError();
}

hipError_t event_global_occupancy(int lds_bytes, int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                        (const void*)event_kernel<true>, kWave,
                                                        (size_t)lds_bytes);
}

}  // namespace mcl
