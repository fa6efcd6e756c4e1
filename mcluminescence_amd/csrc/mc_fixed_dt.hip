// mc_fixed_dt.hip — fixed-dt trap/charge-kinetics Monte Carlo, one wave64 per cluster.
//
// Replaces the run x sim x step loop of simulate.py:460-516 (Box / Physics /
// TLTrapSim: engine.py:44-188, tl_trap_lab.py:24-60).  One workgroup of one
// wavefront owns one cluster; everything the step touches lives in LDS:
//
//   hx, hy, hz [h_cap]   hole coordinates, float32 nm, SoA (removed hole: hx = 3e38)
//   claim      [h_cap]   per-hole lowest firing electron index (conflict rule)
//   ex, ey, ez [e_cap]   electron coordinates, indexed by original electron index
//   est        [e_cap]   compacted live electrons: (nearest hole << 16) | electron index
//   ew         [e_cap]   compacted w = exp(-alpha r_nn), float32
//   ethr       [e_cap]   fire threshold of a candidate in the current step (conflict retries)
//
// Per step n (start-of-step temperature T_n = T0 + T_rate n dt, simulate.py:477-479):
//   scalars   A = s e^{-E_cb/kT}, B_i = b e^{-E_loc_i/kT} in float64 (own exp),
//             x_A = A dt, b_i dt in float32; the threshold of x_A once per step.
//   fire      per live electron: u = Philox(n>>2, e, run, sim|FIRE)[n&3],
//             x = fma(b dt, w, x_A); fire iff u < (uint32)((1 - e^{-x}) 2^32).
//             When fma(b dt, w, x_A) == x_A (tunnelling below one ulp of the
//             band term) the per-step threshold is reused: same function of x,
//             no per-electron transcendental.
//   conflict  ds_min_u32 of the electron index into claim[nn]; the lowest index
//             keeps the hole (src/est_params/functions.py:228-239).
//   remove    winners' slots become dead, their holes get the x sentinel.
//   repair    survivors whose hole died re-search: one wave scan over the holes,
//             (d2 bits, index) min-reduced across the wave = first-index argmin.
//   retry     a conflict loser retries on its new nearest hole within the step,
//             accepted with probability T_new / T_old (one more Philox word), for
//             up to MCL_MAX_ROUNDS rounds; this removes the O(k dt) delay a
//             "loser waits a step" rule would put into the decay curves.
//   compact   stable ballot/mbcnt compaction, only when it drops a 64-slot row.
// Counts go out per bin (bin of step n = n * n_bins / n_steps), zero-padded
// after the last electron recombines, like the reference's padded rows.
#include "mcl_device.h"
#include "mcl_internal.h"
#include "../../include/mcl.h"

namespace mcl {

__device__ __forceinline__ float unit_coord(uint32_t word, double side)
{
    return (float)(((double)word * 0x1p-32) * side);
}

__global__ __launch_bounds__(64) void fixed_dt_kernel(LaunchArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int c = blockIdx.x;
    if (c >= a.n_clusters) return;
    const int lane = (int)threadIdx.x;
    const int run = uniform(a.c_run[c]);
    const int sim = uniform(a.c_sim[c]);
    const RunDesc R = a.runs[run];
    const int n_e0 = uniform(R.n_e0);
    const int n_h = uniform(R.n_h);
    const int n_steps = uniform(R.n_steps);
    const int h_cap = a.h_cap, e_cap = a.e_cap;

    float* hx = (float*)lds_raw;
    float* hy = hx + h_cap;
    float* hz = hy + h_cap;
    uint32_t* claim = (uint32_t*)(hz + h_cap);
    float* ex = (float*)(claim + h_cap);
    float* ey = ex + e_cap;
    float* ez = ey + e_cap;
    uint32_t* est = (uint32_t*)(ez + e_cap);
    float* ew = (float*)(est + e_cap);
    uint32_t* ethr = (uint32_t*)(ew + e_cap);

    const uint32_t k0 = a.key0, k1 = a.key1;
    const uint32_t prun = (uint32_t)R.run_id;
    const uint32_t psim = (uint32_t)sim << 4;

    // ---- geometry (Box.seed, engine.py:124-129) --------------------------------
    const int n_h4 = (n_h + 3) & ~3;
    for (int h = lane; h < n_h4; h += kWave) {
        float x = kDeadHoleX, y = 0.0f, z = 0.0f;
        if (h < n_h) {
            const u32x4 o = philox10(0u, (uint32_t)h, prun, psim | MCL_TAG_GEOM_H, k0, k1);
            x = unit_coord(o.x, R.side_b_nm);
            y = unit_coord(o.y, R.side_b_nm);
            z = unit_coord(o.z, R.side_b_nm);
        }
        hx[h] = x;
        hy[h] = y;
        hz[h] = z;
        claim[h] = 0xFFFFFFFFu;
    }
    for (int e = lane; e < n_e0; e += kWave) {
        const u32x4 o = philox10(0u, (uint32_t)e, prun, psim | MCL_TAG_GEOM_E, k0, k1);
        ex[e] = unit_coord(o.x, R.side_nm);
        ey[e] = unit_coord(o.y, R.side_nm);
        ez[e] = unit_coord(o.z, R.side_nm);
    }
    __syncthreads();

    const float alpha_f = (float)R.alpha_nm;

    // ---- initial nearest hole (Box._rebuild, engine.py:113-119) -----------------
    // Lane per electron; every lane reads the same hole (LDS broadcast), four
    // holes per float4 load, strict < in increasing index = first-index argmin.
    const int rows0 = (n_e0 + kWave - 1) / kWave;
    for (int r = 0; r < rows0; ++r) {
        const int e = r * kWave + lane;
        const bool act = e < n_e0;
        const float x = act ? ex[e] : 0.0f;
        const float y = act ? ey[e] : 0.0f;
        const float z = act ? ez[e] : 0.0f;
        float best = __builtin_inff();
        uint32_t idx = kNoHole;
        for (int h = 0; h < n_h4; h += 4) {
            const float4 X = *(const float4*)(hx + h);
            const float4 Y = *(const float4*)(hy + h);
            const float4 Z = *(const float4*)(hz + h);
#define MCL_NN_STEP(cx, cy, cz, j)                                                   \
            {                                                                        \
                const float dx = cx - x, dy = cy - y, dz = cz - z;                  \
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)); \
                if (d2 < best) { best = d2; idx = (uint32_t)(h + j); }               \
            }
            MCL_NN_STEP(X.x, Y.x, Z.x, 0)
            MCL_NN_STEP(X.y, Y.y, Z.y, 1)
            MCL_NN_STEP(X.z, Y.z, Z.z, 2)
            MCL_NN_STEP(X.w, Y.w, Z.w, 3)
#undef MCL_NN_STEP
        }
        if (act) {
            const float w = idx == kNoHole ? 0.0f : exp_own(-(alpha_f * __builtin_sqrtf(best)));
            est[e] = (idx << 16) | (uint32_t)e;
            ew[e] = w;
        }
    }
    __syncthreads();

    // ---- step loop (simulate.py:475-516) -------------------------------------------
    const int nb = a.n_bins;
    uint32_t* cnt = a.counts + (size_t)c * (size_t)nb;
    const bool two_ch = R.two_channel != 0;
    const bool always2 = R.always_chan2 != 0;
    const uint64_t ret_thr = R.ret_thr;
    const uint32_t c3f = psim | MCL_TAG_FIRE, c3c = psim | MCL_TAG_CHAN;
    const int step_limit = a.dump.enabled ? min(a.dump.max_steps, n_steps) : n_steps;

    int n_live = n_e0, n_slots = n_e0;
    uint64_t lsteps = 0;
    uint64_t dist_evals = (uint64_t)n_e0 * (uint64_t)n_h;
    int steps_run = 0;
    uint32_t bin_acc = 0;
    int bin = 0;
    int64_t bin_rem = 0;
    float xA = 0.0f, b1dt = 0.0f, b2dt = 0.0f;
    FireThr tA{0u, false};

    for (int n = 0; n < step_limit; ++n) {
        if (n_live == 0) break;
        lsteps += (uint64_t)n_live;
        steps_run = n + 1;
        if (n == 0 || R.T_rate != 0.0) {
            const double t = (double)n * R.dt;
            const double T = R.T0_K + R.T_rate * t;
            const double kT = R.k_b * T;
            const double A = R.s * exp_own(-(R.E_cb / kT));
            const double B1 = R.b * exp_own(-(R.E_loc_1 / kT));
            const double B2 = R.b * exp_own(-(R.E_loc_2 / kT));
            xA = (float)(A * R.dt);
            b1dt = (float)(B1 * R.dt);
            b2dt = (float)(B2 * R.dt);
            tA = fire_threshold(xA);
        }
        const uint32_t blk = (uint32_t)n >> 2;
        const int word = n & 3;
        const int rows = (n_slots + kWave - 1) / kWave;

        // round 0: fire test + claims
        uint64_t cand = 0;
        for (int r = 0; r < rows; ++r) {
            const int p = r * kWave + lane;
            const uint32_t st = p < n_slots ? est[p] : kDeadSlot;
            const uint32_t o = st & 0xFFFFu, nn = st >> 16;
            if (nn != kNoHole) {
                const float w = ew[p];
                bool chan2 = always2;
                if (two_ch) {
                    const u32x4 rc = philox10(blk, o, prun, c3c, k0, k1);
                    chan2 = (uint64_t)pick_word(rc, word) < ret_thr;
                }
                const u32x4 rf = philox10(blk, o, prun, c3f, k0, k1);
                const uint32_t u = pick_word(rf, word);
                const float x = __builtin_fmaf(chan2 ? b2dt : b1dt, w, xA);
                FireThr t = tA;
                if (x != xA) t = fire_threshold(x);
                if (t.always || u < t.thr) {
                    atomicMin(&claim[nn], o);
                    ethr[p] = t.always ? 0xFFFFFFFFu : t.thr;
                    cand |= 1ull << r;
                }
            }
        }
        uint64_t rowmask = wave_or_u64(cand);
        uint32_t wins = 0;
        if (rowmask != 0) {
            for (int round = 0;; ++round) {
                if (round > 0) {
                    // claims of this round's candidates (losers accepted for a retry)
                    uint64_t m = rowmask;
                    while (m != 0) {
                        const int r = __builtin_ctzll(m);
                        m &= m - 1;
                        if ((cand >> r) & 1ull) {
                            const uint32_t st = est[r * kWave + lane];
                            atomicMin(&claim[st >> 16], st & 0xFFFFu);
                        }
                    }
                }
                __syncthreads();
                // winners: claim[nn] holds the lowest candidate index for that hole
                uint64_t lose = 0;
                uint64_t m = rowmask;
                while (m != 0) {
                    const int r = __builtin_ctzll(m);
                    m &= m - 1;
                    const int p = r * kWave + lane;
                    const bool is_cand = (cand >> r) & 1ull;
                    bool win = false;
                    uint32_t nn = 0;
                    if (is_cand) {
                        const uint32_t st = est[p];
                        nn = st >> 16;
                        win = claim[nn] == (st & 0xFFFFu);
                        if (!win) lose |= 1ull << r;
                    }
                    if (win) {
                        claim[nn] = 0xFFFFFFFFu;
                        hx[nn] = kDeadHoleX;
                        est[p] = kDeadSlot;
                    }
                    wins += (uint32_t)__popcll(ballot(win));
                }
                __syncthreads();

                // repair: survivors whose nearest hole was removed re-search
                for (int r = 0; r < rows; ++r) {
                    const int p = r * kWave + lane;
                    const uint32_t st = p < n_slots ? est[p] : kDeadSlot;
                    const uint32_t nn = st >> 16;
                    const uint32_t nn_safe = nn != kNoHole ? nn : 0u;
                    const bool need = nn != kNoHole && hx[nn_safe] == kDeadHoleX;
                    uint64_t mm = ballot(need);
                    while (mm != 0) {
                        const int l = __builtin_ctzll(mm);
                        mm &= mm - 1;
                        const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)st, l) & 0xFFFFu;
                        const float x = ex[o], y = ey[o], z = ez[o];
                        uint64_t key = ~0ull;
                        for (int h = lane; h < n_h4; h += kWave) {
                            const float dx = hx[h] - x, dy = hy[h] - y, dz = hz[h] - z;
                            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                            const uint64_t k = ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)h;
                            key = k < key ? k : key;
                        }
                        key = wave_min_u64(key);
                        dist_evals += (uint64_t)n_h;
                        const uint32_t d2bits = (uint32_t)(key >> 32);
                        const bool none = d2bits >= 0x7f800000u;
                        const uint32_t nn_new = none ? kNoHole : (uint32_t)key;
                        const float w =
                            none ? 0.0f : exp_own(-(alpha_f * __builtin_sqrtf(__uint_as_float(d2bits))));
                        if (lane == 0) {
                            est[r * kWave + l] = (nn_new << 16) | o;
                            ew[r * kWave + l] = w;
                        }
                    }
                }
                __syncthreads();
                if (round + 1 == MCL_MAX_ROUNDS) break;

                // a loser retries on its new nearest hole in the same step, accepted
                // with probability T_new / T_old (thinning of the draw it already won)
                const uint64_t lmask = wave_or_u64(lose);
                if (lmask == 0) break;
                cand = 0;
                const uint32_t c3r = psim | (MCL_TAG_RETRY0 + (uint32_t)round);
                m = lmask;
                while (m != 0) {
                    const int r = __builtin_ctzll(m);
                    m &= m - 1;
                    const int p = r * kWave + lane;
                    if ((lose >> r) & 1ull) {
                        const uint32_t st = est[p];
                        const uint32_t o = st & 0xFFFFu, nn = st >> 16;
                        if (nn != kNoHole) {
                            bool chan2 = always2;
                            if (two_ch) {
                                const u32x4 rc = philox10(blk, o, prun, c3c, k0, k1);
                                chan2 = (uint64_t)pick_word(rc, word) < ret_thr;
                            }
                            const float x = __builtin_fmaf(chan2 ? b2dt : b1dt, ew[p], xA);
                            FireThr t = tA;
                            if (x != xA) t = fire_threshold(x);
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
                rowmask = wave_or_u64(cand);
                if (rowmask == 0) break;
            }
            n_live -= (int)wins;

            // lazy stable compaction when a 64-slot row can be dropped
            if ((n_live + kWave - 1) / kWave < rows) {
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
                __syncthreads();
            }
        }
        bin_acc += wins;
        bin_rem += nb;
        if (bin_rem >= n_steps) {
            bin_rem -= n_steps;
            if (lane == 0) cnt[bin] = bin_acc;
            bin_acc = 0;
            ++bin;
        }
    }
    if (bin < nb) {
        if (lane == 0) cnt[bin] = bin_acc;
        ++bin;
    }
    for (int b = bin + lane; b < nb; b += kWave) cnt[b] = 0u;
    if (a.live_steps != nullptr && lane == 0) a.live_steps[c] = lsteps;
    if (a.work != nullptr && lane == 0) {
        a.work[4 * (size_t)c] = lsteps;
        a.work[4 * (size_t)c + 1] = dist_evals;
        a.work[4 * (size_t)c + 2] = (uint64_t)(n_e0 - n_live);
        a.work[4 * (size_t)c + 3] = (uint64_t)steps_run;
    }

    if (a.dump.enabled) {
        const StateDump& d = a.dump;
        __syncthreads();
        for (int e = lane; e < n_e0; e += kWave) {
            d.e_live[e] = 0;
            d.nn[e] = kNoHole;
            d.w[e] = 0.0f;
            d.e_xyz[3 * e] = ex[e];
            d.e_xyz[3 * e + 1] = ey[e];
            d.e_xyz[3 * e + 2] = ez[e];
        }
        __syncthreads();
        for (int p = lane; p < n_slots; p += kWave) {
            const uint32_t st = est[p];
            if (st != kDeadSlot) {
                const uint32_t o = st & 0xFFFFu;
                d.e_live[o] = 1;
                d.nn[o] = st >> 16;
                d.w[o] = ew[p];
            }
        }
        for (int h = lane; h < n_h; h += kWave) {
            d.h_alive[h] = hx[h] != kDeadHoleX;
            d.h_xyz[3 * h] = hx[h];
            d.h_xyz[3 * h + 1] = hy[h];
            d.h_xyz[3 * h + 2] = hz[h];
        }
    }
}

hipError_t prepare_fixed_dt(int lds_bytes)
{
    return hipFuncSetAttribute((const void*)fixed_dt_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

hipError_t launch_fixed_dt(const LaunchArgs& a, int lds_bytes, hipStream_t stream)
{
    hipLaunchKernelGGL(fixed_dt_kernel, dim3((unsigned)a.n_clusters), dim3(kWave),
                       (unsigned)lds_bytes, stream, a);
    return hipGetLastError();
}

}  // namespace mcl
