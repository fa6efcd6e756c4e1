// mc_lab.hip — the lab experiments of TLTrapSim.TL_lab / ISO_lab
// (tl_trap_lab.py:65-179): an event-mode Monte Carlo with dose filling
// (_filling_time, tl_trap_lab.py:53-60; Box.add_electron, engine.py:133-152) that
// records the trapped fraction n_e / N_e at observation times.  One wave64 per
// (experiment, sim).  Lab clusters are small (N_e ~ 100), so electrons and holes
// keep explicit float32 coordinates in LDS in creation order — the order of the
// reference's arrays, where add_electron appends and remove_pair deletes — and a
// nearest-hole search is a full wave scan (first index on ties).
//
// Step (t = time so far, float64):
//   T = (T_start + 273.15) + T_rate t; A = s e^{-E_cb/kT}, B_c = b e^{-E_loc_c/kT};
//   k_e = fma(B_ch, w_e, A) per live electron (channel as in the other kernels);
//   K summed by 64-electron row scans added in order (as the event kernel).
//   One Philox block (i, 0xFFFFFFFF, run, sim|FIRE): E_t, U_w, E_f as in the event
//   kernel.  lambda = 1e-20 if D = 0 or n_e = int(N_e), else (D/D0)(int(N_e) - n_e);
//   dt_fill = (1 / lambda) E_f; dt_rec = E_t / K (dt_fill without electrons, +inf
//   when K = 0); dt = min(dt_rec, dt_fill), no cap (tl_trap_lab.py:86-89,142-144).
//   t += dt.  dt == dt_fill: a fill — electron f of the run at Philox
//   (f, 0xFFFFFFFE, run, sim|GEOM_E) in the core cube and a hole at
//   (f, 0xFFFFFFFE, run, sim|GEOM_H) in the boundary cube.  The new electron's
//   nearest hole is searched among the holes before the new one and the other
//   electrons keep theirs (engine.py:141-152, the reference's stale update), or,
//   with exact_fill, among all holes with every electron compared against the new
//   hole.  Otherwise a recombination: the electron whose running sum first
//   exceeds U_w K goes with its nearest hole; electrons that had that hole search
//   again (engine.py:154-175).
//   After the step every observation time <= t not yet passed records n_e / N_e
//   (tl_trap_lab.py:163-172); the run ends when all are recorded.
//
// Options of the legacy lab loops (src/est_params/functions.py:270-434):
//   MCL_LAB_SHARED_CHANNEL  one channel draw per step for every electron
//                           (lifetime_fading, functions.py:115-135), word i & 3 of
//                           Philox (i >> 2, 0xFFFFFFFD, run, sim|CHAN);
//   MCL_LAB_FIRST_ELECTRON  the recombining electron is the first one in creation
//                           order that has a hole (recomber as the lab loops call it,
//                           functions.py:241-261: its np.where over a (1, n_e) mask
//                           returns row indices, all 0);
//   MCL_LAB_RETRAP          after a recombination, with probability Retrap (word w
//                           of the step block), one more pair with exact nearest
//                           holes (functions.py:328-331).
#include "mc_common.h"

namespace mcl {

namespace {

constexpr uint32_t kFillCtr = 0xFFFFFFFEu;
constexpr uint32_t kSharedChanCtr = 0xFFFFFFFDu;


__device__ __forceinline__ double lab_exp1(uint32_t u)
{
    return -log_own(((double)u + 0.5) * 0x1p-32);
}

// First-index nearest live hole among slots [0, n) for the point (x, y, z):
// returns (d2 bits << 32 | slot), d2 = +inf (no live hole) as 0x7f800000.
__device__ __forceinline__ uint64_t lab_nearest(const float4* hole, int n, float x, float y,
                                                float z, int lane)
{
    uint64_t key = ~0ull;
    for (int s = lane; s < n; s += kWave) {
        const float4 h = hole[s];
        const float dx = h.x - x, dy = h.y - y, dz = h.z - z;
        const float d2 = h.x == kDeadHoleX ? __builtin_inff()
                                           : __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const uint64_t k = ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)s;
        key = k < key ? k : key;
    }
    return wave_min_u64(key);
}

// The same for every lane's own point at once: each lane scans all n slots (the
// lanes read the same slot together, one LDS broadcast), the same key per
// (point, slot) as lab_nearest, so the same minimum.
__device__ __forceinline__ uint64_t lab_nearest_own(const float4* hole, int n, float x, float y,
                                                    float z)
{
    uint64_t key = ~0ull;
    for (int s = 0; s < n; ++s) {
        const float4 h = hole[s];
        const float dx = h.x - x, dy = h.y - y, dz = h.z - z;
        const float d2 = h.x == kDeadHoleX ? __builtin_inff()
                                           : __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const uint64_t k = ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)s;
        key = k < key ? k : key;
    }
    return key;
}

}  // namespace

__device__ __forceinline__ void run_cluster_lab(const LabArgs& a, const int c,
                                                unsigned char* lds_raw)
{
    const int lane = (int)threadIdx.x;
    const int run = uniform(a.c_run[c]);
    const int sim = uniform(a.c_sim[c]);
    const RunDesc R = a.runs[run];
    const int h_cap = a.h_cap, e_cap = a.e_cap;
    float4* hole = (float4*)lds_raw;                  // [h_cap] {x, y, z, -}, creation order
    float4* elec = hole + h_cap;                      // [e_cap] {x, y, z, w}, creation order
    uint32_t* enn = (uint32_t*)(elec + e_cap);        // [e_cap] nearest hole slot
    double* rtot = (double*)(enn + e_cap);            // [e_cap / 64] row sums
    uint32_t* remap = (uint32_t*)(rtot + e_cap / kWave);   // [h_cap] slot map of a hole compaction

    const uint32_t k0 = a.key0, k1 = a.key1;
    const uint32_t prun = (uint32_t)R.run_id;
    const uint32_t psim = (uint32_t)sim << 4;
    const float alpha_f = (float)R.alpha_nm;
    const int N_int = (int)R.N_e;

    // ---- seed (Box.seed, engine.py:124-129) --------------------------------------
    int n_e = uniform(R.n_e0);
    int n_hs = uniform(R.n_h);                        // hole slots in use (live or dead)
    for (int h = lane; h < n_hs; h += kWave) {
        const u32x4 o = philox10(0u, (uint32_t)h, prun, psim | MCL_TAG_GEOM_H, k0, k1);
        hole[h] = make_float4(unit_coord(o.x, R.side_b_nm), unit_coord(o.y, R.side_b_nm),
                              unit_coord(o.z, R.side_b_nm), 0.0f);
    }
    for (int e = lane; e < n_e; e += kWave) {
        const u32x4 o = philox10(0u, (uint32_t)e, prun, psim | MCL_TAG_GEOM_E, k0, k1);
        elec[e] = make_float4(unit_coord(o.x, R.side_nm), unit_coord(o.y, R.side_nm),
                              unit_coord(o.z, R.side_nm), 0.0f);
    }
    __syncthreads();
    // initial nearest holes (Box._rebuild, engine.py:113-119), one electron a lane
    for (int e = lane; e < n_e; e += kWave) {
        const float4 p = elec[e];
        const uint64_t key = lab_nearest_own(hole, n_hs, p.x, p.y, p.z);
        uint32_t slot;
        float w;
        nn_from_key(key, alpha_f, slot, w);
        enn[e] = slot;
        elec[e].w = w;
    }
    __syncthreads();

    const bool two_ch = R.two_channel != 0;
    const bool always2 = R.always_chan2 != 0;
    const bool exact_fill = (a.options & MCL_LAB_EXACT_FILL) != 0;
    const bool shared_ch = (a.options & MCL_LAB_SHARED_CHANNEL) != 0;
    const bool first_e = (a.options & MCL_LAB_FIRST_ELECTRON) != 0;
    const bool retrap = (a.options & MCL_LAB_RETRAP) != 0;
    const uint64_t ret_thr = R.ret_thr;
    const uint32_t c3f = psim | MCL_TAG_FIRE, c3c = psim | MCL_TAG_CHAN;
    const bool b1_is_a = R.b == R.s && R.E_loc_1 == R.E_cb;
    const bool b2_is_b1 = R.E_loc_2 == R.E_loc_1;
    const double* obs = a.obs + a.obs_offset[run];
    const int n_obs = a.obs_count[run];
    double* out = a.ratio + (size_t)c * (size_t)a.max_obs;

    double t = 0.0;
    int i = 0, k_obs = 0, fills = 0;
    int32_t status = 0;
    uint64_t lsteps = 0, recomb = 0;
    // The step draws depend on the step index only: lane j holds those of step
    // sb_base + j (E_t, U_w, E_f and the retrap word, each computed as the step
    // would), 64 steps in one parallel evaluation; a step reads its own with
    // scalar reads.
    int sb_base = -kWave;
    uint32_t l_uw = 0u, l_rw = 0u;
    double l_et = 0.0, l_ef = 0.0;
    while (k_obs < n_obs) {
        if ((int64_t)i >= a.max_steps) {
            status = -1;
            break;
        }
        lsteps += (uint64_t)n_e;
        const double T = R.T0_K + R.T_rate * t;
        const double kT = R.k_b * T;
        const double A = R.s * exp_own(-(R.E_cb / kT));
        const double B1 = b1_is_a ? A : R.b * exp_own(-(R.E_loc_1 / kT));
        const double B2 = b2_is_b1 ? B1 : R.b * exp_own(-(R.E_loc_2 / kT));
        const uint32_t blk = (uint32_t)i >> 2;
        const int word = i & 3;
        bool ch_step = always2;
        if (two_ch && shared_ch)
            ch_step = (uint64_t)pick_word(philox10(blk, kSharedChanCtr, prun, c3c, k0, k1), word) <
                      ret_thr;
        // k_e of slot p, 0 for slots past the live ones and electrons with no hole;
        // branch-free: every slot of a row lies within e_cap, so the reads are in
        // bounds, and the selects drop what the slots past n_e hold
        auto rate_of = [&](int p) -> double {
            const uint32_t nn = enn[p];
            const float w = elec[p].w;
            bool ch = ch_step;
            if (two_ch && !shared_ch)
                ch = (uint64_t)pick_word(philox10(blk, (uint32_t)p, prun, c3c, k0, k1), word) <
                     ret_thr;
            const double k = __builtin_fma(ch ? B2 : B1, (double)w, A);
            return p < n_e && nn != kNoHole ? k : 0.0;
        };
        // one electron-hole pair (Box.add_electron); false when the cluster is full
        auto add_pair = [&](bool exact) -> bool {
            if (n_e >= e_cap) return false;
            if (n_hs >= h_cap) {
                // stable compaction of the hole slots (creation order kept)
                int base = 0;
                for (int s0 = 0; s0 < n_hs; s0 += kWave) {
                    const int s = s0 + lane;
                    const bool live = s < n_hs && hole[s].x != kDeadHoleX;
                    const float4 hv = s < n_hs ? hole[s] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const uint64_t m = ballot(live);
                    const int q = base + lanes_below(m);
                    if (s < n_hs) remap[s] = live ? (uint32_t)q : kNoHole;
                    __syncthreads();
                    if (live) hole[q] = hv;
                    __syncthreads();
                    base += __popcll(m);
                }
                for (int e = lane; e < n_e; e += kWave)
                    if (enn[e] != kNoHole) enn[e] = remap[enn[e]];
                n_hs = base;
                __syncthreads();
                if (n_hs >= h_cap) return false;
            }
            const uint32_t f = (uint32_t)fills;
            const u32x4 oe = philox10(f, kFillCtr, prun, psim | MCL_TAG_GEOM_E, k0, k1);
            const u32x4 oh = philox10(f, kFillCtr, prun, psim | MCL_TAG_GEOM_H, k0, k1);
            const float ex = unit_coord(oe.x, R.side_nm), ey = unit_coord(oe.y, R.side_nm),
                        ez = unit_coord(oe.z, R.side_nm);
            const float4 hn = make_float4(unit_coord(oh.x, R.side_b_nm),
                                          unit_coord(oh.y, R.side_b_nm),
                                          unit_coord(oh.z, R.side_b_nm), 0.0f);
            const int hs = n_hs;
            __syncthreads();
            if (lane == 0) hole[hs] = hn;
            __syncthreads();
            // the new electron's nearest hole: before the new hole (reference) or all
            const uint64_t key = lab_nearest(hole, exact ? hs + 1 : hs, ex, ey, ez, lane);
            uint32_t slot;
            float w;
            nn_from_key(key, alpha_f, slot, w);
            if (exact) {
                // existing electrons: the new hole wins only when strictly closer
                for (int e = lane; e < n_e; e += kWave) {
                    const float4 p = elec[e];
                    const uint32_t s0 = enn[e];
                    const float dx = hn.x - p.x, dy = hn.y - p.y, dz = hn.z - p.z;
                    const float d2n = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    float d2o = __builtin_inff();
                    if (s0 != kNoHole) {
                        const float4 h = hole[s0];
                        const float ox = h.x - p.x, oy = h.y - p.y, oz = h.z - p.z;
                        d2o = __builtin_fmaf(oz, oz, __builtin_fmaf(oy, oy, ox * ox));
                    }
                    if (d2n < d2o) {
                        enn[e] = (uint32_t)hs;
                        elec[e].w = w_of(__float_as_uint(d2n), alpha_f);
                    }
                }
            }
            __syncthreads();
            if (lane == 0) {
                elec[n_e] = make_float4(ex, ey, ez, w);
                enn[n_e] = slot;
            }
            __syncthreads();
            n_e += 1;
            n_hs += 1;
            fills += 1;
            return true;
        };
        const int rows = (n_e + kWave - 1) / kWave;
        double K = 0.0;
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
        __syncthreads();
        if (i - sb_base >= kWave) {
            const u32x4 o = philox10((uint32_t)(i + lane), 0xFFFFFFFFu, prun, c3f, k0, k1);
            l_et = lab_exp1(o.x);
            l_uw = o.y;
            l_ef = lab_exp1(o.z);
            l_rw = o.w;
            sb_base = i;
        }
        const int js = i - sb_base;
        const double lam = (R.D == 0.0 || n_e == N_int) ? 1e-20
                                                         : (R.D / R.D0) * (double)(N_int - n_e);
        const double dt_fill = (1.0 / lam) * lane_value_d(l_ef, js);
        double dt_rec = dt_fill;
        if (n_e > 0) dt_rec = K > 0.0 ? lane_value_d(l_et, js) / K : __builtin_inf();
        const double dt = fmin(dt_rec, dt_fill);
        t += dt;
        if (dt == dt_fill) {
            // ---- filling: one electron-hole pair (Box.add_electron) ------------
            if (!add_pair(exact_fill)) {
                status = -2;
                break;
            }
        } else if (dt == dt_rec) {
            // ---- recombination (Box.remove_pair) --------------------------------
            int pw = -1;
            if (first_e) {
                for (int r = 0; r < rows && pw < 0; ++r) {
                    const int p = r * kWave + lane;
                    const uint64_t m = ballot(p < n_e && enn[p] != kNoHole);
                    if (m != 0) pw = r * kWave + (int)__builtin_ctzll(m);
                }
            } else {
                const double target = (double)lane_value(l_uw, js) * 0x1p-32 * K;
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
            }
            if (pw < 0) {
                for (int r = rows - 1; r >= 0 && pw < 0; --r) {
                    const uint64_t m = ballot(rate_of(r * kWave + lane) > 0.0);
                    if (m != 0) pw = r * kWave + 63 - (int)__builtin_clzll(m);
                }
            }
            pw = uniform(pw);
            const uint32_t hw = enn[pw];
            __syncthreads();
            if (lane == 0) hole[hw].x = kDeadHoleX;
            // electrons after pw move down one slot (np.delete keeps the order)
            for (int r0 = 0; r0 < n_e; r0 += kWave) {
                const int p = r0 + lane;
                const bool mv = p > pw && p < n_e;
                const float4 v = mv ? elec[p] : make_float4(0.f, 0.f, 0.f, 0.f);
                const uint32_t nv = mv ? enn[p] : 0u;
                __syncthreads();
                if (mv) {
                    elec[p - 1] = v;
                    enn[p - 1] = nv;
                }
                __syncthreads();
            }
            n_e -= 1;
            recomb += 1;
            // the electrons whose nearest hole was hw, found row by row by a ballot, each
            // searched by the whole wave (the holes do not change meanwhile, so the
            // order of the searches does not matter)
            for (int r0 = 0; r0 < n_e; r0 += kWave) {
                uint64_t m = ballot(r0 + lane < n_e && enn[r0 + lane] == hw);
                while (m != 0) {
                    const int e = r0 + (int)__builtin_ctzll(m);
                    m &= m - 1;
                    const float4 p = elec[e];
                    const uint64_t key = lab_nearest(hole, n_hs, p.x, p.y, p.z, lane);
                    if (lane == 0) {
                        uint32_t slot;
                        float w;
                        nn_from_key(key, alpha_f, slot, w);
                        enn[e] = slot;
                        elec[e].w = w;
                    }
                }
            }
            __syncthreads();
            // legacy retrapping: a new pair after the recombination
            if (retrap && (uint64_t)lane_value(l_rw, js) < R.retrap_thr && !add_pair(true)) {
                status = -2;
                break;
            }
        }
        ++i;
        while (k_obs < n_obs && t >= obs[k_obs]) {
            if (lane == 0) out[k_obs] = (double)n_e / R.N_e;
            ++k_obs;
        }
    }
    for (int k = k_obs + lane; k < a.max_obs; k += kWave) out[k] = 0.0;
    if (lane == 0) {
        a.status[c] = status;
        if (a.work != nullptr) {
            a.work[4 * (size_t)c] = lsteps;
            a.work[4 * (size_t)c + 1] = (uint64_t)fills;
            a.work[4 * (size_t)c + 2] = recomb;
            a.work[4 * (size_t)c + 3] = (uint64_t)i;
        }
    }
}

__global__ __launch_bounds__(64) void lab_kernel(LabArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if ((int)blockIdx.x < a.n_clusters) run_cluster_lab(a, (int)blockIdx.x, lds_raw);
}

int lab_lds_bytes(int h_cap, int e_cap)
{
    return 16 * h_cap + 16 * e_cap + 4 * e_cap + 8 * (e_cap / kWaveHost) + 4 * h_cap;
}

hipError_t prepare_lab(int lds_bytes)
{
    return hipFuncSetAttribute((const void*)lab_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lds_bytes);
}

hipError_t launch_lab(const LabArgs& a, int lds_bytes, hipStream_t stream)
{
    hipLaunchKernelGGL(lab_kernel, dim3((unsigned)a.n_clusters), dim3(kWave), (unsigned)lds_bytes,
                       stream, a);
    return hipGetLastError();
}

}  // namespace mcl
