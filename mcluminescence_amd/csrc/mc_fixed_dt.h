// mc_fixed_dt.h — fixed-dt trap/charge-kinetics Monte Carlo, one workgroup per cluster:
// the per-cluster body shared by the LDS kernels (mc_fixed_dt.hip) and the
// global-memory kernels for clusters beyond one workgroup's LDS (mc_fixed_dt_global.hip).
//
// Replaces the run x sim x step loop of simulate.py:36-92 (Box / Physics /
// TLTrapSim: engine.py:44-188, tl_trap_lab.py:24-60).  One workgroup of NW
// wavefronts (1, 2, 4 or 8) owns one cluster; everything the step touches lives in LDS:
//
//   hole       [h_cap]   16-byte records {x, y, z, meta}, float32 nm, sorted by
//                        grid cell; meta = (claiming electron << 16) | original
//                        hole index; a removed hole keeps its slot with x = 3e38
//   cstart     [g_cap]   first sorted slot of each cell, cstart[G^3] = n_h
//   est        [e_cap]   compacted live electrons: (nearest hole slot << 16) | electron index
//   ew         [e_cap]   compacted w = exp(-alpha r_nn), float32
//   ethr       [e_cap]   fire threshold of a candidate in the current step (conflict
//                        retries); the draw of an electron queued for a new search
//   rq         [e_cap]   electrons whose nearest hole must be searched again: (slot << 16) | e
//   xch        [2 NW + 1] (NW > 1) per-wave values exchanged between the waves, and the
//                        persistent launch's ticket
//
// A cluster too large for that (GM: the global-memory form) keeps hole and
// cstart in its workgroup's slice of a global scratch buffer (slice blockIdx.x,
// so its launches are persistent: one slice per resident workgroup, reused
// cluster after cluster); xch opens the LDS block, followed by est, ew, ethr and
// rq when 16 e_cap bytes fit there (e_lds), otherwise those follow cstart in the
// slice.  The code is the same; the workgroup barriers order global memory as
// they order LDS (every wave of a workgroup runs on one CU, behind one L1), and
// the queue fences of several waves widen to workgroup scope, which waits for
// the wave's global stores.  Only the slots holes take within a cell can differ
// from the LDS form (global atomics need not serve the lanes of one instruction
// in lane order); no result depends on them: every search keys on
// (distance, original hole index).
//
// With NW > 1 the 64-slot rows of est/ew are split into NW contiguous ranges, one
// per wave.  Every per-electron decision is the one-wave kernel's (same draws,
// same thresholds, same lowest-index conflict rule, same stable compaction
// order), so the counts are identical for every NW; the waves meet at workgroup
// barriers where one wave's results reach the others: after the fire pass (any
// candidate?), after each conflict round's winners (wins, any loser?), at each
// retry round's claims, and in a compaction (per-wave keep counts; the rows are
// rewritten into the spare est/ew pair, rq and ethr, whose roles then swap).
// Each wave searches the queue of its own rows in its own part of rq.
//
// Electron coordinates are not stored: they are a pure function of the
// Philox counter (0, e, run, sim|GEOM_E) and are regenerated when an
// electron's nearest hole has to be searched again.
//
// Nearest holes come from a uniform G^3 cell grid over the hole cube: a search
// scans the 3x3x3 cells around the electron and accepts the minimum of
// (d2 bits, original hole index) only if it is closer than every hole outside
// that block can be (guaranteed face distance, with a margin far above float32
// rounding); otherwise the whole wave scans every hole for that electron.  The
// answer is therefore always the first-index argmin of the full scan
// (engine.py:118-119, 173-175); the grid only removes work.  All electrons that
// need a search in a step are queued and searched lane-parallel, so a step pays
// the latency of one search, not one per recombination; each lane fetches the
// holes of a cell range eight 16-byte records at a time, so one LDS latency
// covers a chunk.
//
// Per step n (start-of-step temperature T_n = T0 + T_rate n dt, simulate.py:53-54):
//   scalars   A = s e^{-E_cb/kT}, B_i = b e^{-E_loc_i/kT} in float64 (own exp),
//             x_A = A dt, b_i dt in float32; the threshold of x_A once per step.
//   fire      per live electron: u = Philox(n>>2, e, run, sim|FIRE)[n&3],
//             x = fma(b dt, w, x_A); fire iff u < (uint32)((1 - e^{-x}) 2^32).
//             When fma(b dt, w, x_A) == x_A (tunnelling below one ulp of the
//             band term) the per-step threshold is reused: same function of x,
//             no per-electron transcendental.  Two rows per iteration keep two
//             independent Philox chains in flight.
//   conflict  ds_min_u32 of the electron index into the hole's claim field; the
//             lowest index keeps the hole (src/est_params/functions.py:228-239).
//   remove    winners' slots become dead, their holes get the x sentinel.
//   repair    an electron whose hole was removed keeps its old w until a step in
//             which its draw would fire against that old w: only then is it
//             searched again (before its fire test; a conflict loser at once).
//             Exact: the new nearest hole is no closer, so w and x = k dt can only
//             fall, and the threshold of x is non-decreasing in x (checked for
//             every float x by the oracle), so a draw that fails against the old
//             w fails against the new one.
//   retry     a conflict loser retries on its new nearest hole within the step,
//             accepted with probability T_new / T_old (one more Philox word), for
//             up to MCL_MAX_ROUNDS rounds; this removes the O(k dt) delay a
//             "loser waits a step" rule would put into the decay curves.
//   compact   stable ballot/mbcnt compaction, only when it drops a 64-slot row.
// Counts go out per bin (bin of step n = n * n_bins / n_steps), zero-padded
// after the last electron recombines, like the reference's padded rows.
#pragma once

#include "mc_common.h"

#include <algorithm>

namespace mcl {

constexpr int kCachePairs = 4;    // row pairs whose Philox blocks stay in registers
template <int V>
struct CacheSlot {
    static constexpr int value = V;
};

// Ordering between the steps of one wave that write and then read the search
// queue: the workgroup barrier with one wave (where it is also the compiler's
// fence), a fence with several (each wave owns its queue; a workgroup barrier
// there would not be reached by every wave): wavefront scope when the queue is
// in LDS, whose instructions one wave executes in issue order; workgroup scope
// in the global-memory form, so the wave's global stores have completed before
// its lanes read them.
template <int NW, bool GM>
__device__ __forceinline__ void queue_sync()
{
    if constexpr (NW == 1) {
        __syncthreads();
    } else if constexpr (GM) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Simulate cluster c from its geometry to its last step with NW wavefronts, all
// state in lds_raw (GM: holes and cell starts, and the electron arrays unless
// a.e_lds, in the workgroup's scratch slice).  Every branch that holds a
// workgroup barrier depends only on values all waves share (n, n_live, n_slots,
// and results of the exchanges).  FW: the build with the fixed-threshold path
// (isothermal one-channel runs keep fire thresholds in their w slots); without
// it every run computes its thresholds from w, and the code of that path is not
// in the kernel.  QW: the build with the quiet window (after quiet steps, the rest
// of a Philox block tested at once).  Every build gives the same counts for every
// run.
template <int NW, bool GM, bool FW, bool QW>
__device__ __forceinline__ void run_cluster(const LaunchArgs& a, const int c,
                                            unsigned char* lds_raw)
{
    const int tid = (int)threadIdx.x;
    const int lane = NW == 1 ? tid : (tid & (kWave - 1));
    const int wv = NW == 1 ? 0 : uniform(tid >> 6);
    const int run = uniform(a.c_run[c]);
    const int sim = uniform(a.c_sim[c]);
    const RunDesc R = a.runs[run];
    const int n_e0 = uniform(R.n_e0);
    const int n_h = uniform(R.n_h);
    const int n_steps = uniform(R.n_steps);
    const int h_cap = a.h_cap, e_cap = a.e_cap, g_cap = a.g_cap;

    unsigned char* hbase;      // hole, cstart
    unsigned char* ebase;      // est, ew, ethr, rq
    uint32_t* xbase;           // xch
    if constexpr (GM) {
        unsigned char* const slice = a.gscratch + (size_t)blockIdx.x * (size_t)a.gstride;
        hbase = slice;
        xbase = (uint32_t*)lds_raw;
        ebase = a.e_lds ? lds_raw + fixed_dt_waves_lds(NW)
                        : slice + 16 * (size_t)h_cap + 4 * (size_t)g_cap;
    } else {
        hbase = lds_raw;
        ebase = lds_raw + 16 * h_cap + 4 * g_cap;
        xbase = (uint32_t*)(ebase + 16 * e_cap);
    }
    float4* hole = (float4*)hbase;
    float* hole_f = (float*)hbase;                   // word view: x at 4 s, meta at 4 s + 3
    uint32_t* hole_u = (uint32_t*)hbase;
    uint32_t* cstart = (uint32_t*)(hole + h_cap);
    uint32_t* est = (uint32_t*)ebase;
    float* ew = (float*)(est + e_cap);
    uint32_t* ethr = (uint32_t*)(ew + e_cap);
    uint32_t* rq = ethr + e_cap;
    uint32_t* const xch = xbase;

    const uint32_t k0 = a.key0, k1 = a.key1;
    const uint32_t prun = (uint32_t)R.run_id;
    const uint32_t psim = (uint32_t)sim << 4;
    const ElectronGen egen{prun, psim | MCL_TAG_GEOM_E, k0, k1, R.side_nm};

    const bool two_ch = R.two_channel != 0;
    const bool always2 = R.always_chan2 != 0;
    // single-trap parameter sets share one Boltzmann factor between the channels
    const bool b1_is_a = R.b == R.s && R.E_loc_1 == R.E_cb;
    const bool b2_is_b1 = R.E_loc_2 == R.E_loc_1;
    // Step scalars of step m: start-of-step temperature, Boltzmann factors in
    // float64, their float32 products with dt, the threshold of the band term.
    struct StepScalars {
        float xA, b1dt, b2dt;
        FireThr tA;
    };
    auto scalars_at = [&](int m) {
        const double t = (double)m * R.dt;
        const double T = R.T0_K + R.T_rate * t;
        const double kT = R.k_b * T;
        const double A = R.s * exp_own(-(R.E_cb / kT));
        const double B1 = b1_is_a ? A : R.b * exp_own(-(R.E_loc_1 / kT));
        const double B2 = b2_is_b1 ? B1 : R.b * exp_own(-(R.E_loc_2 / kT));
        StepScalars sc;
        sc.xA = (float)(A * R.dt);
        sc.b1dt = (float)(B1 * R.dt);
        sc.b2dt = (float)(B2 * R.dt);
        sc.tA = fire_threshold(sc.xA);
        return sc;
    };
    const StepScalars sc0 = scalars_at(0);
    // An isothermal run with one localisation channel keeps each electron's fire
    // threshold, constant while its w is, in the w slot (state dumps keep w).
    const bool fixedw = FW && R.T_rate == 0.0 && !two_ch && a.dump.enabled == 0;
    const WSlot wslot{fixedw, always2 ? sc0.b2dt : sc0.b1dt, sc0.xA};

    uint64_t dist_evals = 0;
    const Grid grid = setup_cluster_t<NW>(R, egen, hole, cstart, est, ew, prun, psim, k0, k1, tid,
                                          dist_evals, wslot);
    const float alpha_f = (float)R.alpha_nm;

    // ---- step loop (simulate.py:51-92) -------------------------------------------
    const int nb = a.n_bins > 0 ? a.n_bins : n_steps;   // n_bins == 0: one bin per step
    uint32_t* cnt = a.counts + (size_t)c * (size_t)a.row_len;
    const uint64_t ret_thr = R.ret_thr;
    const uint32_t c3f = psim | MCL_TAG_FIRE, c3c = psim | MCL_TAG_CHAN;
    const int step_limit = a.dump.enabled ? min(a.dump.max_steps, n_steps) : n_steps;

    int n_live = n_e0, n_slots = n_e0;
    uint64_t lsteps = 0;
    int steps_run = 0;
    uint32_t bin_acc = 0;
    int bin = 0;
    int64_t bin_rem = 0;
    float xA = 0.0f, b1dt = 0.0f, b2dt = 0.0f;
    FireThr tA{0u, false};

    auto hole_dead = [&](uint32_t nn) {
        return nn != kNoHole && hole_f[4 * (nn != kNoHole ? nn : 0u)] == kDeadHoleX;
    };
    // Claim hole nn for electron o: the claim field keeps the lowest index.
    auto claim = [&](uint32_t nn, uint32_t o) {
        atomicMin(&hole_u[4 * nn + 3], (o << 16) | (hole_u[4 * nn + 3] & 0xFFFFu));
    };
    // Fire threshold of the electron whose w slot holds v, in channel ch (2 when
    // set): the one kept in the slot, or that of x = k dt = fma(b dt, w, x_A).
    auto slot_thr = [&](float v, bool ch) -> FireThr {
        if (fixedw) return slot_threshold(v);
        const float x = __builtin_fmaf(ch ? b2dt : b1dt, v, xA);
        return x != xA ? fire_threshold(x) : tA;
    };
    // Test of one live electron (slot p) with draw u and threshold t; candidates
    // claim their hole and remember their threshold for a conflict retry.
    auto fire_tail = [&](uint32_t st, uint32_t o, int p, uint32_t u, FireThr t, int rr,
                         uint64_t& cand) {
        const uint32_t nn = st >> 16;
        if (nn == kNoHole) return;
        if (t.always || u < t.thr) {
            claim(nn, o);
            ethr[p] = t.always ? 0xFFFFFFFFu : t.thr;
            cand |= 1ull << rr;
        }
    };
    // Pass-A test of one live electron (slot p, draw u, threshold t of its current
    // w): nothing unless the draw fires; then one 16-byte read of its hole gives
    // both the removed-hole marker and the claim field.  A removed hole queues the
    // electron for a new search (its draw kept in ethr; the test against the old w
    // bounds the one against the new); a live one is claimed.  Returns queued.
    auto test_fire = [&](uint32_t st, uint32_t o, int p, uint32_t u, FireThr t, int rr,
                         uint64_t& cand) -> bool {
        const uint32_t nn = st >> 16;
        if (nn == kNoHole) return false;
        if (!(t.always || u < t.thr)) return false;
        const float4 h = hole[nn];
        if (h.x == kDeadHoleX) {
            ethr[p] = u;
            return true;
        }
        atomicMin(&hole_u[4 * nn + 3], (o << 16) | (__float_as_uint(h.w) & 0xFFFFu));
        ethr[p] = t.always ? 0xFFFFFFFFu : t.thr;
        cand |= 1ull << rr;
        return false;
    };
    // Search again for the queued electrons q[0..nq): one electron at a time
    // with the whole wave while the queue is short, lane-parallel otherwise.
    auto search_queue = [&](const uint32_t* q, int nq) {
        search_queue_nn(grid, egen, est, ew, q, nq, a.wave_search_max, alpha_f, lane, dist_evals,
                        wslot);
    };
    // Per-wave values between the waves (NW > 1): lane 0 of each wave publishes
    // v, one workgroup barrier, then every wave reads all NW values.  Consecutive
    // exchanges alternate between two slots of xch and each holds a barrier, so a
    // slot is rewritten only after every wave has read it.
    int xph = 0;
    auto exchange = [&](uint32_t v) -> const uint32_t* {
        uint32_t* slot = xch + (xph & 1) * NW;
        if (lane == 0) slot[wv] = v;
        __syncthreads();
        ++xph;
        return slot;
    };
    auto any_wave = [&](bool v) -> bool {
        if constexpr (NW == 1) {
            return v;
        } else {
            const uint32_t* x = exchange(v ? 1u : 0u);
            uint32_t r = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) r |= x[w];
            return uniform(r) != 0u;
        }
    };

    // Look-ahead for few live electrons: lane j tests step n + j for every live
    // electron with exactly the draw and threshold that step would use (a stale
    // electron against its old w, as in pass A).  Returns how many steps from n on
    // no electron can fire (at most 64, up to step_limit); found tells whether
    // some step of the window may fire.  Such quiet steps change nothing but
    // the live-step and bin tallies.
    auto quiet_steps = [&](int n0, bool& found) -> int {
        const int len = min(kWave, step_limit - n0);
        const int m = n0 + lane;
        StepScalars sc{xA, b1dt, b2dt, tA};
        if (R.T_rate != 0.0) sc = scalars_at(m);
        const uint32_t mblk = (uint32_t)m >> 2;
        const int mword = m & 3;
        bool fire = false;
        for (int p = 0; p < n_slots; ++p) {
            const uint32_t st = est[p];                  // one address: a broadcast read
            if (st == kDeadSlot || (st >> 16) == kNoHole) continue;
            const uint32_t o = st & 0xFFFFu;
            const float w = ew[p];
            const uint32_t u = pick_word(philox10(mblk, o, prun, c3f, k0, k1), mword);
            bool ch = always2;
            if (two_ch)
                ch = (uint64_t)pick_word(philox10(mblk, o, prun, c3c, k0, k1), mword) < ret_thr;
            FireThr t;
            if (fixedw) {
                t = slot_threshold(w);
            } else {
                const float x = __builtin_fmaf(ch ? sc.b2dt : sc.b1dt, w, sc.xA);
                t = x != sc.xA ? fire_threshold(x) : sc.tA;
            }
            fire = fire || t.always || u < t.thr;
        }
        const uint64_t fm = ballot(fire && lane < len);
        found = fm != 0;
        return found ? (int)__builtin_ctzll(fm) : len;
    };
    // Bin tally at the end of a step with `wins` recombinations.
    auto finish_step = [&](uint32_t wins) {
        bin_acc += wins;
        bin_rem += nb;
        if (bin_rem >= n_steps) {
            bin_rem -= n_steps;
            if (tid == 0) cnt[bin] = bin_acc;
            bin_acc = 0;
            ++bin;
        }
    };

    xA = sc0.xA;
    b1dt = sc0.b1dt;
    b2dt = sc0.b2dt;
    tA = sc0.tA;
    // Philox blocks of the first 2 kCachePairs rows, kept in registers: steps
    // n .. n+3 share the counter (n >> 2, e, run, sim|tag), so a block computed in
    // one of them serves the other three.  cache_pairs marks the row pairs whose
    // blocks are those of block cache_blk for the electrons now in their slots.
    u32x4 cfa[kCachePairs], cfb[kCachePairs];   // fire blocks of rows 2q, 2q + 1
    uint32_t cch[kCachePairs];                   // channel bits: row 2q word k at bit k,
                                                 // row 2q + 1 at bit 4 + k
    uint32_t cache_pairs = 0u;
    int cache_blk = -1;
    float l_xA = 0.0f, l_b1dt = 0.0f, l_b2dt = 0.0f;   // step scalars of step sc_base + lane
    uint32_t l_thr = 0u, l_always = 0u;
    int sc_base = -kWave;
    // The step scalars of step m >= 1 of a ramp, from lane m - sc_base (the 64 steps
    // from sc_base are evaluated in parallel, refilled from m when m leaves them).
    auto ramp_scalars = [&](int m, int last) -> StepScalars {
        if (m < sc_base || last - sc_base >= kWave) {
            const StepScalars sc = scalars_at(m + lane);
            l_xA = sc.xA;
            l_b1dt = sc.b1dt;
            l_b2dt = sc.b2dt;
            l_thr = sc.tA.thr;
            l_always = sc.tA.always ? 1u : 0u;
            sc_base = m;
        }
        const int j = m - sc_base;
        StepScalars sc;
        sc.xA = lane_value(l_xA, j);
        sc.b1dt = lane_value(l_b1dt, j);
        sc.b2dt = lane_value(l_b2dt, j);
        sc.tA.thr = lane_value(l_thr, j);
        sc.tA.always = lane_value(l_always, j) != 0u;
        return sc;
    };
    // Quiet window: the steps n0 .. n0 + W - 1 (n0 >= 1, within one Philox block, so
    // one block per electron holds all their draws) tested at once, each live
    // electron's draw of each step against that step's threshold exactly as pass A
    // tests it (a stale electron against its old w).  Returns how many steps from
    // n0 no electron of the cluster fires in (W: none); those steps change nothing
    // but the tallies, and the next step that runs finds its cached rows' blocks in
    // the register cache.
    auto quiet_window = [&](int n0, int W) -> int {
        const uint32_t wblk = (uint32_t)n0 >> 2;
        const int wd = n0 & 3;
        const int wrows = (n_slots + kWave - 1) / kWave;
        const int per = (wrows + NW - 1) / NW;      // rows of a wave, at most (all waves)
        // only rows whose blocks the register cache keeps (the next step reads them
        // there); with more rows a wave, the steps run one by one (the same choice
        // in every wave: a window holds an exchange)
        if (per > 2 * kCachePairs) return 0;
        int w_lo = 0, wnr = wrows, w_end = n_slots;
        if constexpr (NW > 1) {
            w_lo = min(wrows, wv * per);
            wnr = min(wrows, w_lo + per) - w_lo;
            w_end = min(n_slots, (w_lo + wnr) * kWave);
        }
        if (R.T_rate != 0.0) (void)ramp_scalars(n0, n0 + W - 1);   // the window's steps in the lanes
        // step scalars of step n0 + k
        auto scal = [&](int k) -> StepScalars {
            if (R.T_rate == 0.0) return StepScalars{xA, b1dt, b2dt, tA};
            return ramp_scalars(n0 + k, n0 + k);
        };
        if ((int)wblk != cache_blk) {
            cache_blk = (int)wblk;
            cache_pairs = 0u;
        }
        // the first step in which an electron of this wave fires (W: none); a pair
        // tests only the steps before the first one an earlier pair found
        int lim = W;
        auto window_pair = [&](int r, auto slot_tag) {
            constexpr int q = decltype(slot_tag)::value;
            const int p0 = (w_lo + r) * kWave + lane, p1 = p0 + kWave;
            const uint32_t st0 = p0 < w_end ? est[p0] : kDeadSlot;
            const uint32_t st1 = p1 < w_end ? est[p1] : kDeadSlot;
            const float v0 = p0 < w_end ? ew[p0] : 0.0f, v1 = p1 < w_end ? ew[p1] : 0.0f;
            const uint32_t o0 = st0 & 0xFFFFu, o1 = st1 & 0xFFFFu;
            u32x4 fa, fb;
            uint32_t cb = 0u;
            bool hit = false;
            if constexpr (q >= 0) hit = ((cache_pairs >> q) & 1u) != 0u;
            if (hit) {
                if constexpr (q >= 0) {
                    fa = cfa[q];
                    fb = cfb[q];
                    cb = cch[q];
                }
            } else {
                fa = philox10(wblk, o0, prun, c3f, k0, k1);
                fb = philox10(wblk, o1, prun, c3f, k0, k1);
                if (two_ch) {
                    const u32x4 ca = philox10(wblk, o0, prun, c3c, k0, k1);
                    const u32x4 cc = philox10(wblk, o1, prun, c3c, k0, k1);
                    cb = (uint32_t)((uint64_t)ca.x < ret_thr) | (uint32_t)((uint64_t)ca.y < ret_thr) << 1 |
                         (uint32_t)((uint64_t)ca.z < ret_thr) << 2 | (uint32_t)((uint64_t)ca.w < ret_thr) << 3 |
                         (uint32_t)((uint64_t)cc.x < ret_thr) << 4 | (uint32_t)((uint64_t)cc.y < ret_thr) << 5 |
                         (uint32_t)((uint64_t)cc.z < ret_thr) << 6 | (uint32_t)((uint64_t)cc.w < ret_thr) << 7;
                }
                if constexpr (q >= 0) {
                    cfa[q] = fa;
                    cfb[q] = fb;
                    cch[q] = cb;
                    cache_pairs |= 1u << q;
                }
            }
            for (int k = 0; k < lim; ++k) {
                const StepScalars sc = scal(k);
                const int d = wd + k;
                const bool ch0 = two_ch ? ((cb >> d) & 1u) != 0u : always2;
                const bool ch1 = two_ch ? ((cb >> (4 + d)) & 1u) != 0u : always2;
                FireThr t0, t1;
                if (fixedw) {
                    t0 = slot_threshold(v0);
                    t1 = slot_threshold(v1);
                } else {
                    const float x0 = __builtin_fmaf(ch0 ? sc.b2dt : sc.b1dt, v0, sc.xA);
                    const float x1 = __builtin_fmaf(ch1 ? sc.b2dt : sc.b1dt, v1, sc.xA);
                    t0 = x0 != sc.xA ? fire_threshold(x0) : sc.tA;
                    t1 = x1 != sc.xA ? fire_threshold(x1) : sc.tA;
                }
                const bool f0 = (st0 >> 16) != kNoHole && (t0.always || pick_word(fa, d) < t0.thr);
                const bool f1 = (st1 >> 16) != kNoHole && (t1.always || pick_word(fb, d) < t1.thr);
                if (ballot(f0 || f1) != 0) {
                    lim = k;
                    break;
                }
            }
        };
        if (NW == 1 || wnr > 0) window_pair(0, CacheSlot<0>{});
        if (wnr > 2) window_pair(2, CacheSlot<1>{});
        if (wnr > 4) window_pair(4, CacheSlot<2>{});
        if (wnr > 6) window_pair(6, CacheSlot<3>{});
        if constexpr (NW > 1) {
            const uint32_t* x = exchange((uint32_t)lim);
            uint32_t m = (uint32_t)W;
#pragma unroll
            for (int w = 0; w < NW; ++w) m = x[w] < m ? x[w] : m;
            lim = (int)uniform(m);
        }
        return lim;
    };
    int quiet_run = 0;           // steps since the last one with a candidate in some wave
    int quiet_need = 2;          // quiet steps before a window; doubled after a window
                                 // whose first step fires, back to 2 after a whole one
    int n = 0;
    while (n < step_limit && n_live > 0) {
        if (n_live <= a.lookahead_max) {
            bool found = false;
            const int f = quiet_steps(n, found);
            for (int k = 0; k < f; ++k) {
                lsteps += (uint64_t)n_live;
                finish_step(0u);
            }
            n += f;
            steps_run = n;
            if (!found) continue;
        } else if (QW && quiet_run >= quiet_need && (n & 3) <= 2) {
            // a quiet stretch: the rest of the block at once
            const int W = min(4 - (n & 3), step_limit - n);
            const int f = quiet_window(n, W);
            for (int k = 0; k < f; ++k) {
                lsteps += (uint64_t)n_live;
                finish_step(0u);
            }
            n += f;
            steps_run = n;
            quiet_run += f;
            if (f == 0) quiet_need = min(2 * quiet_need, 64);
            if (f == W) {
                quiet_need = 2;
                continue;
            }
        }
        lsteps += (uint64_t)n_live;
        steps_run = n + 1;
        if (n > 0 && R.T_rate != 0.0) {
            // a ramp: lane j holds the scalars of step sc_base + j (64 steps in one
            // parallel evaluation instead of 64 serial ones)
            const StepScalars sc = ramp_scalars(n, n);
            xA = sc.xA;
            b1dt = sc.b1dt;
            b2dt = sc.b2dt;
            tA = sc.tA;
        }
        const uint32_t blk = (uint32_t)n >> 2;
        const int word = n & 3;
        const int rows = (n_slots + kWave - 1) / kWave;
        // this wave's rows: [r_lo, r_lo + nr), slots below slot_end (all rows with NW = 1)
        int r_lo = 0, nr = rows, slot_end = n_slots;
        if constexpr (NW > 1) {
            const int per = (rows + NW - 1) / NW;
            r_lo = min(rows, wv * per);
            nr = min(rows, r_lo + per) - r_lo;
            slot_end = min(n_slots, (r_lo + nr) * kWave);
        }
        uint32_t* const rq_w = rq + r_lo * kWave;     // this wave's part of the queue
        const float bmax = fmaxf(b1dt, b2dt);

        // Pass A, two rows per iteration: an electron whose nearest hole was removed
        // and whose draw would fire against its old w is queued for a new search
        // (its test waits for it); one whose draw fails even against the old w stays
        // as it is; every other live electron takes its fire test now.  Deferring
        // the search to here changes nothing: no hole disappears in between.
        uint64_t cand = 0, stale = 0, stale_ch = 0;
        int nq = 0;
        if ((int)blk != cache_blk) {
            cache_blk = (int)blk;
            cache_pairs = 0u;
        }
        auto pass_a_pair = [&](int r, auto slot_tag) {      // r: the wave's row index
            constexpr int q = decltype(slot_tag)::value;   // register cache slot, -1: none
            const int p0 = (r_lo + r) * kWave + lane, p1 = p0 + kWave;
            const uint32_t st0 = p0 < slot_end ? est[p0] : kDeadSlot;
            const uint32_t st1 = p1 < slot_end ? est[p1] : kDeadSlot;
            const float w0 = p0 < slot_end ? ew[p0] : 0.0f, w1 = p1 < slot_end ? ew[p1] : 0.0f;
            const uint32_t o0 = st0 & 0xFFFFu, o1 = st1 & 0xFFFFu;
            // Below x = 2^-32 (in either channel) the threshold is 0: no draw can
            // fire, and a stale electron keeps waiting.  Rows whose lanes all sit
            // there (early in a heating ramp) skip their draws.  (A slot that keeps
            // the threshold holds 0 exactly then.)
            const bool live0 = fixedw ? __float_as_uint(w0) != 0u
                                      : __builtin_fmaf(bmax, w0, xA) >= 0x1p-32f;
            const bool live1 = fixedw ? __float_as_uint(w1) != 0u
                                      : __builtin_fmaf(bmax, w1, xA) >= 0x1p-32f;
            if (ballot(live0 || live1) == 0) return;
            uint32_t u0 = 0u, u1 = 0u;
            bool ch0 = always2, ch1 = always2;
            bool hit = false;
            if constexpr (q >= 0) hit = ((cache_pairs >> q) & 1u) != 0u;
            if (hit) {
                if constexpr (q >= 0) {
                    u0 = pick_word(cfa[q], word);
                    u1 = pick_word(cfb[q], word);
                    if (two_ch) {
                        ch0 = ((cch[q] >> word) & 1u) != 0u;
                        ch1 = ((cch[q] >> (4 + word)) & 1u) != 0u;
                    }
                }
            } else {
                const u32x4 fa = philox10(blk, o0, prun, c3f, k0, k1);
                const u32x4 fb = philox10(blk, o1, prun, c3f, k0, k1);
                u0 = pick_word(fa, word);
                u1 = pick_word(fb, word);
                uint32_t cb = 0u;
                if (two_ch) {
                    const u32x4 ca = philox10(blk, o0, prun, c3c, k0, k1);
                    const u32x4 cc = philox10(blk, o1, prun, c3c, k0, k1);
                    cb = (uint32_t)((uint64_t)ca.x < ret_thr) | (uint32_t)((uint64_t)ca.y < ret_thr) << 1 |
                         (uint32_t)((uint64_t)ca.z < ret_thr) << 2 | (uint32_t)((uint64_t)ca.w < ret_thr) << 3 |
                         (uint32_t)((uint64_t)cc.x < ret_thr) << 4 | (uint32_t)((uint64_t)cc.y < ret_thr) << 5 |
                         (uint32_t)((uint64_t)cc.z < ret_thr) << 6 | (uint32_t)((uint64_t)cc.w < ret_thr) << 7;
                    ch0 = ((cb >> word) & 1u) != 0u;
                    ch1 = ((cb >> (4 + word)) & 1u) != 0u;
                }
                if constexpr (q >= 0) {
                    cfa[q] = fa;
                    cfb[q] = fb;
                    cch[q] = cb;
                    cache_pairs |= 1u << q;
                }
            }
            const bool s0 = test_fire(st0, o0, p0, u0, slot_thr(w0, ch0), r, cand);
            const bool s1 = test_fire(st1, o1, p1, u1, slot_thr(w1, ch1), r + 1, cand);
            const uint64_t m0 = ballot(s0);
            if (s0) rq_w[nq + lanes_below(m0)] = ((uint32_t)p0 << 16) | o0;
            nq += __popcll(m0);
            const uint64_t m1 = ballot(s1);
            if (s1) rq_w[nq + lanes_below(m1)] = ((uint32_t)p1 << 16) | o1;
            nq += __popcll(m1);
            stale |= ((uint64_t)s0 << r) | ((uint64_t)s1 << (r + 1));
            stale_ch |= ((uint64_t)ch0 << r) | ((uint64_t)ch1 << (r + 1));
        };
        if (NW == 1 || nr > 0) pass_a_pair(0, CacheSlot<0>{});
        if (nr > 2) pass_a_pair(2, CacheSlot<1>{});
        if (nr > 4) pass_a_pair(4, CacheSlot<2>{});
        if (nr > 6) pass_a_pair(6, CacheSlot<3>{});
        for (int r = 2 * kCachePairs; r < nr; r += 2) pass_a_pair(r, CacheSlot<-1>{});
        if (nq > 0) {
            queue_sync<NW, GM>();
            search_queue(rq_w, nq);
            queue_sync<NW, GM>();
            uint64_t m = rows_any(stale, nr);
            while (m != 0) {
                const int r = __builtin_ctzll(m);
                m &= m - 1;
                const int p = (r_lo + r) * kWave + lane;
                if ((stale >> r) & 1ull) {
                    // the draw and channel of pass A, the new w
                    const uint32_t st = est[p];
                    const bool ch = (stale_ch >> r) & 1ull;
                    fire_tail(st, st & 0xFFFFu, p, ethr[p], slot_thr(ew[p], ch), r, cand);
                }
            }
        }
        uint64_t rowmask = rows_any(cand, nr);
        uint32_t wins = 0;          // this wave's winners in this step
        uint32_t step_wins = 0;     // the cluster's
        const bool any_cand = any_wave(rowmask != 0);
        quiet_run = any_cand ? 0 : quiet_run + 1;
        if (any_cand) {
            for (int round = 0;; ++round) {
                if (round > 0) {
                    // claims of this round's candidates (losers accepted for a retry)
                    uint64_t m = rowmask;
                    while (m != 0) {
                        const int r = __builtin_ctzll(m);
                        m &= m - 1;
                        if ((cand >> r) & 1ull) {
                            const uint32_t st = est[(r_lo + r) * kWave + lane];
                            claim(st >> 16, st & 0xFFFFu);
                        }
                    }
                    if constexpr (NW > 1) {
                        if (!any_wave(rowmask != 0)) break;   // also the claims' barrier
                    }
                }
                if constexpr (NW == 1) __syncthreads();
                // winners: the claim field holds the lowest candidate index for that hole
                uint64_t lose = 0;
                uint64_t m = rowmask;
                while (m != 0) {
                    const int r = __builtin_ctzll(m);
                    m &= m - 1;
                    const int p = (r_lo + r) * kWave + lane;
                    const bool is_cand = (cand >> r) & 1ull;
                    bool win = false;
                    uint32_t nn = 0, meta = 0;
                    if (is_cand) {
                        const uint32_t st = est[p];
                        nn = st >> 16;
                        meta = hole_u[4 * nn + 3];
                        win = (meta >> 16) == (st & 0xFFFFu);
                        if (!win) lose |= 1ull << r;
                    }
                    if (win) {
                        hole_u[4 * nn + 3] = 0xFFFF0000u | (meta & 0xFFFFu);
                        hole_f[4 * nn] = kDeadHoleX;
                        est[p] = kDeadSlot;
                    }
                    wins += (uint32_t)__popcll(ballot(win));
                }
                uint64_t lrows;
                if constexpr (NW == 1) {
                    __syncthreads();
                    if (round + 1 == MCL_MAX_ROUNDS) break;
                    lrows = rows_any(lose, rows);
                    if (lrows == 0) break;
                } else {
                    // the removed holes are visible to every wave after this barrier
                    lrows = rows_any(lose, nr);
                    const uint32_t* x = exchange(wins | (lrows != 0 ? 0x80000000u : 0u));
                    uint32_t t = 0, l = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        t += x[w] & 0x7FFFFFFFu;
                        l |= x[w] >> 31;
                    }
                    step_wins = uniform(t);
                    if (round + 1 == MCL_MAX_ROUNDS) break;
                    if (uniform(l) == 0u) break;
                }

                // a loser's hole is gone: search again now, then it retries in the same
                // step, accepted with probability T_new / T_old (thinning of the draw
                // it already won).  Other electrons that lost their hole are searched
                // again at the start of the next step.
                int nl = 0;
                m = lrows;
                while (m != 0) {
                    const int r = __builtin_ctzll(m);
                    m &= m - 1;
                    const int p = (r_lo + r) * kWave + lane;
                    const bool l = (lose >> r) & 1ull;
                    const uint64_t mm = ballot(l);
                    if (l) rq_w[nl + lanes_below(mm)] = ((uint32_t)p << 16) | (est[p] & 0xFFFFu);
                    nl += __popcll(mm);
                }
                queue_sync<NW, GM>();
                search_queue(rq_w, nl);
                queue_sync<NW, GM>();
                cand = 0;
                const uint32_t c3r = psim | (MCL_TAG_RETRY0 + (uint32_t)round);
                m = lrows;
                while (m != 0) {
                    const int r = __builtin_ctzll(m);
                    m &= m - 1;
                    const int p = (r_lo + r) * kWave + lane;
                    if ((lose >> r) & 1ull) {
                        const uint32_t st = est[p];
                        const uint32_t o = st & 0xFFFFu, nn = st >> 16;
                        if (nn != kNoHole) {
                            bool chan2 = always2;
                            if (two_ch) {
                                const u32x4 rc = philox10(blk, o, prun, c3c, k0, k1);
                                chan2 = (uint64_t)pick_word(rc, word) < ret_thr;
                            }
                            const FireThr t = slot_thr(ew[p], chan2);
                            const uint64_t Tn = t.always ? 0x100000000ull : (uint64_t)t.thr;
                            const uint32_t eo = ethr[p];
                            const uint64_t To = eo == 0xFFFFFFFFu ? 0x100000000ull : (uint64_t)eo;
                            bool acc = Tn >= To;
                            if (!acc) {
                                const u32x4 rr = philox10(blk, o, prun, c3r, k0, k1);
                                acc = (uint64_t)pick_word(rr, word) * To < (Tn << 32);
                            }
                            if (acc) {
                                ethr[p] = t.always ? 0xFFFFFFFFu : t.thr;
                                cand |= 1ull << r;
                            }
                        }
                    }
                }
                rowmask = rows_any(cand, nr);
                if constexpr (NW == 1) {
                    if (rowmask == 0) break;
                }
            }
            if constexpr (NW == 1) step_wins = wins;
            n_live -= (int)step_wins;

            // lazy stable compaction when a 64-slot row can be dropped (an electron
            // waiting for a new search moves with its stale nearest hole)
            if ((n_live + kWave - 1) / kWave < rows) {
                if constexpr (NW == 1) {
                    int base = 0;
                    for (int r = 0; r < rows; ++r) {
                        const int p = r * kWave + lane;
                        const uint32_t st = p < n_slots ? est[p] : kDeadSlot;
                        const float w = p < n_slots ? ew[p] : 0.0f;
                        const bool keep = st != kDeadSlot;
                        const uint64_t km = ballot(keep);
                        if (keep) {
                            const int q = base + lanes_below(km);
                            est[q] = st;
                            ew[q] = w;
                        }
                        base += __popcll(km);
                    }
                    n_slots = base;
                } else {
                    // per-wave keep counts give each wave's first new slot; the rows
                    // are rewritten into the spare pair (rq, ethr), which becomes est/ew
                    uint32_t kept = 0;
                    for (int r = 0; r < nr; ++r) {
                        const int p = (r_lo + r) * kWave + lane;
                        const uint32_t st = p < slot_end ? est[p] : kDeadSlot;
                        kept += (uint32_t)__popcll(ballot(st != kDeadSlot));
                    }
                    const uint32_t* x = exchange(kept);
                    uint32_t before = 0, total = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        before += w < wv ? x[w] : 0u;
                        total += x[w];
                    }
                    int base = (int)uniform(before);
                    uint32_t* const n_est = rq;
                    float* const n_ew = (float*)ethr;
                    for (int r = 0; r < nr; ++r) {
                        const int p = (r_lo + r) * kWave + lane;
                        const uint32_t st = p < slot_end ? est[p] : kDeadSlot;
                        const float w = p < slot_end ? ew[p] : 0.0f;
                        const bool keep = st != kDeadSlot;
                        const uint64_t km = ballot(keep);
                        if (keep) {
                            const int q = base + lanes_below(km);
                            n_est[q] = st;
                            n_ew[q] = w;
                        }
                        base += __popcll(km);
                    }
                    rq = est;
                    ethr = (uint32_t*)ew;
                    est = n_est;
                    ew = n_ew;
                    n_slots = (int)uniform(total);
                }
                cache_pairs = 0u;                   // electrons moved to other slots
                __syncthreads();
            }
        }
        finish_step(step_wins);
        ++n;
    }
    if (bin < nb) {
        if (tid == 0) cnt[bin] = bin_acc;
        ++bin;
    }
    for (int b = bin + tid; b < a.row_len; b += NW * kWave) cnt[b] = 0u;
    if (a.live_steps != nullptr && tid == 0) a.live_steps[c] = lsteps;

    // (state dumps are one-wave launches only)
    if (NW == 1 && a.dump.enabled) {
        // searches deferred to the next step happen now, so the dump shows every
        // survivor's current nearest hole
        __syncthreads();
        int nq = 0;
        for (int r = 0; r * kWave < n_slots; ++r) {
            const int p = r * kWave + lane;
            const uint32_t st = p < n_slots ? est[p] : kDeadSlot;
            const bool s = hole_dead(st >> 16);
            const uint64_t m = ballot(s);
            if (s) rq[nq + lanes_below(m)] = ((uint32_t)p << 16) | (st & 0xFFFFu);
            nq += __popcll(m);
        }
        __syncthreads();
        search_queue(rq, nq);
        __syncthreads();
    }

    // lanes counted their own evaluations in the lane-private searches
    uint64_t evals_total = dist_evals;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)evals_total, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(evals_total >> 32), off);
        evals_total += ((uint64_t)hi << 32) | lo;
    }
    if constexpr (NW > 1) {
        const uint32_t* xl = exchange((uint32_t)evals_total);
        uint64_t lo_sum = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) lo_sum += xl[w];
        const uint32_t* xh = exchange((uint32_t)(evals_total >> 32));
        uint64_t hi_sum = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) hi_sum += xh[w];
        evals_total = lo_sum + (hi_sum << 32);
    }
    if (a.work != nullptr && tid == 0) {
        a.work[4 * (size_t)c] = lsteps;
        a.work[4 * (size_t)c + 1] = evals_total;
        a.work[4 * (size_t)c + 2] = (uint64_t)(n_e0 - n_live);
        a.work[4 * (size_t)c + 3] = (uint64_t)steps_run;
    }

    if (NW == 1 && a.dump.enabled) {
        const StateDump& d = a.dump;
        for (int e = lane; e < n_e0; e += kWave) {
            float x, y, z;
            egen.xyz((uint32_t)e, x, y, z);
            d.e_live[e] = 0;
            d.nn[e] = kNoHole;
            d.w[e] = 0.0f;
            d.e_xyz[3 * e] = x;
            d.e_xyz[3 * e + 1] = y;
            d.e_xyz[3 * e + 2] = z;
        }
        __syncthreads();
        for (int p = lane; p < n_slots; p += kWave) {
            const uint32_t st = est[p];
            const uint32_t o = st & 0xFFFFu, slot = st >> 16;
            if (st != kDeadSlot && o < (uint32_t)n_e0) {
                d.e_live[o] = 1;
                d.nn[o] = slot == kNoHole ? kNoHole : (hole_u[4 * slot + 3] & 0xFFFFu);
                d.w[o] = ew[p];
            }
        }
        for (int s = lane; s < n_h; s += kWave) {
            const float4 hv = hole[s];
            const uint32_t h = __float_as_uint(hv.w) & 0xFFFFu;
            if (h >= (uint32_t)n_h) continue;
            d.h_alive[h] = hv.x != kDeadHoleX;
            d.h_xyz[3 * h] = hv.x;
            d.h_xyz[3 * h + 1] = hv.y;
            d.h_xyz[3 * h + 2] = hv.z;
        }
    }
}

// One workgroup per cluster (queue == null: cluster order[blockIdx.x], or
// blockIdx.x when order is null), or persistent workgroups that take clusters
// from order[] (null: 0 .. n_order-1) through a global atomic ticket until the
// list is exhausted.  Every workgroup leaves the loop when the ticket passes
// n_order.  The global-memory form always runs persistent.
template <bool GM, bool FW, bool QW>
__global__ __launch_bounds__(64, GM ? 2 : 1) void fixed_dt_kernel(LaunchArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (a.queue == nullptr) {
        const int i = blockIdx.x;
        if (i < a.n_order) run_cluster<1, GM, FW, QW>(a, a.order != nullptr ? a.order[i] : i, lds_raw);
        return;
    }
    for (;;) {
        int i = 0;
        if (threadIdx.x == 0) i = (int)atomicAdd(a.queue, 1u);
        i = __shfl(i, 0);
        if (i >= a.n_order) break;
        run_cluster<1, GM, FW, QW>(a, a.order != nullptr ? a.order[i] : i, lds_raw);
        __syncthreads();
    }
}

// The same with NW wavefronts per cluster (at least two waves per SIMD, so at
// most 256 VGPRs a wave, as the one-wave kernel).  The persistent ticket goes
// through LDS (the word after the exchange slots), read by every wave between
// two barriers, so all waves take the same cluster and leave the loop together.
template <int NW, bool GM, bool FW, bool QW>
__global__ __launch_bounds__(NW * 64, 2) void fixed_dt_mw_kernel(LaunchArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (a.queue == nullptr) {
        const int i = blockIdx.x;
        if (i < a.n_order) run_cluster<NW, GM, FW, QW>(a, a.order != nullptr ? a.order[i] : i, lds_raw);
        return;
    }
    uint32_t* const ticket =
        (GM ? (uint32_t*)lds_raw
            : (uint32_t*)(lds_raw + 16 * a.h_cap + 4 * a.g_cap + 16 * a.e_cap)) + 2 * NW;
    for (;;) {
        if (threadIdx.x == 0) *ticket = atomicAdd(a.queue, 1u);
        __syncthreads();
        const int i = (int)uniform(*ticket);
        __syncthreads();
        if (i >= a.n_order) break;
        run_cluster<NW, GM, FW, QW>(a, a.order != nullptr ? a.order[i] : i, lds_raw);
        __syncthreads();
    }
}

template <bool GM, bool FW, bool QW>
const void* fixed_dt_fn_t(int waves)
{
    return waves == 8 ? (const void*)fixed_dt_mw_kernel<8, GM, FW, QW>
         : waves == 4 ? (const void*)fixed_dt_mw_kernel<4, GM, FW, QW>
         : waves == 2 ? (const void*)fixed_dt_mw_kernel<2, GM, FW, QW>
                      : (const void*)fixed_dt_kernel<GM, FW, QW>;
}

// The build of the NW-wave kernel with the quiet window or not (QW) and with the
// fixed-threshold path or not (fw).
template <bool GM, bool QW>
const void* fixed_dt_fn(int waves, bool fw)
{
    return fw ? fixed_dt_fn_t<GM, true, QW>(waves) : fixed_dt_fn_t<GM, false, QW>(waves);
}

template <bool GM, bool FW, bool QW>
void launch_fixed_dt_w(const LaunchArgs& a, dim3 g, unsigned lds, int waves, hipStream_t stream)
{
    if (waves == 8)
        hipLaunchKernelGGL((fixed_dt_mw_kernel<8, GM, FW, QW>), g, dim3(8 * kWave), lds, stream, a);
    else if (waves == 4)
        hipLaunchKernelGGL((fixed_dt_mw_kernel<4, GM, FW, QW>), g, dim3(4 * kWave), lds, stream, a);
    else if (waves == 2)
        hipLaunchKernelGGL((fixed_dt_mw_kernel<2, GM, FW, QW>), g, dim3(2 * kWave), lds, stream, a);
    else
        hipLaunchKernelGGL((fixed_dt_kernel<GM, FW, QW>), g, dim3(kWave), lds, stream, a);
}

// The builds with the quiet window (QW) or without are instantiated in separate
// sources (mc_fixed_dt*.hip without it, mc_fixed_dt*_window.hip with it), which
// the launch functions choose between by a.quiet_window.
template <bool GM, bool QW>
hipError_t launch_fixed_dt_t(const LaunchArgs& a, int grid, unsigned lds, int waves,
                             hipStream_t stream)
{
    const dim3 g((unsigned)grid);
    if (a.fixed_thr != 0)
        launch_fixed_dt_w<GM, true, QW>(a, g, lds, waves, stream);
    else
        launch_fixed_dt_w<GM, false, QW>(a, g, lds, waves, stream);
    return hipGetLastError();
}

// Resident workgroups per CU of the NW-wave kernel with this much dynamic LDS:
// the smaller count of its two fixed-threshold builds.
template <bool GM, bool QW>
hipError_t fixed_dt_occupancy_t(size_t lds, int waves, int* blocks_per_cu)
{
    int b0 = 0, b1 = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b0, fixed_dt_fn<GM, QW>(waves, false),
                                                                waves * kWave, lds);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, fixed_dt_fn<GM, QW>(waves, true),
                                                         waves * kWave, lds);
    *blocks_per_cu = b0 < b1 ? b0 : b1;
    return e;
}

// Dynamic-LDS limits of the builds of one form (GM) and window choice (QW).
template <bool GM, bool QW>
hipError_t prepare_fixed_dt_t(int bytes_one_wave, int max_lds)
{
    hipError_t e = hipSuccess;
    for (int w : {1, 2, 4, 8}) {
        for (bool fw : {false, true}) {
            int b = bytes_one_wave + fixed_dt_waves_lds(w);
            b = b < max_lds ? b : max_lds;
            if (e == hipSuccess)
                e = hipFuncSetAttribute(fixed_dt_fn<GM, QW>(w, fw),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, b);
        }
    }
    return e;
}

}  // namespace mcl
