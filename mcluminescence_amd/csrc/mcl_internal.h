// mcl_internal.h — host/device shared descriptors of libmcl (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcl {

// Per-run constants of the fixed-dt kernel, derived on the host from
// mcl_run_params (geometry of tl_trap_lab.py:33-39 and engine.py:124-129).
struct RunDesc {
    double side_nm;     // core cube side [nm]
    double side_b_nm;   // hole cube side bf * d [nm]
    double alpha_nm;    // alpha [1/nm]
    double s, b, E_cb, E_loc_1, E_loc_2, k_b;
    double T0_K;        // T_start + 273.15
    double T_rate;      // K/s
    double dt;          // s
    uint64_t ret_thr;   // floor(Retrap * 2^32): channel 2 iff u < ret_thr
    int32_t n_e0, n_h, n_steps;
    int32_t two_channel;   // draw the localisation channel per electron-step
    int32_t always_chan2;  // Retrap >= 1: every draw picks E_loc_2
    int32_t run_id;        // Philox counter word 2 (the run's index in its sweep)
    int32_t grid;          // G: cells per axis of the hole grid (G^3 cells)
    float cell_nm;         // cell edge = side_b_nm / G
    float inv_cell;        // G / side_b_nm
    int32_t pad_;
    // event mode (the reference loop, simulate.py:46-92)
    double T_start_C;      // T_start [degC]: T = T_start + T_rate t + 273.15
    double N_e;            // trap count: electron_ratio = n_e / N_e
    double D, D0;          // dose rate [Gy/s] and D0 of the filling rate
    double duration;       // [s] the loop runs while t <= duration
    uint64_t retrap_thr;   // legacy retrapping: a new pair iff u < retrap_thr (2^32: always)
};

// Optional dump of one cluster's kernel state (tests / inspection).
struct StateDump {
    float* e_xyz;       // [n_e0][3]
    float* h_xyz;       // [n_h][3]
    uint8_t* e_live;    // [n_e0]
    uint8_t* h_alive;   // [n_h]
    uint32_t* nn;       // [n_e0]
    float* w;           // [n_e0]
    int32_t max_steps;  // stop after this many steps and dump
    int32_t enabled;
};

struct LaunchArgs {
    const RunDesc* runs;
    const int32_t* c_run;   // [n_clusters] run index of each cluster
    const int32_t* c_sim;   // [n_clusters] sim index within its run
    int32_t n_clusters;
    const int32_t* order;   // clusters this launch simulates (null: 0 .. n_order-1)
    int32_t n_order;
    uint32_t* queue;        // persistent launch: atomic ticket into order (null: one per block)
    uint32_t key0, key1;    // Philox key = seed
    int32_t n_bins;         // bins per cluster; 0 = one bin per step
    int32_t row_len;        // counts row stride: n_bins, or the longest n_steps
    uint32_t* counts;       // [n_clusters][row_len]
    uint64_t* live_steps;   // [n_clusters] or null
    uint64_t* work;         // [n_clusters][4] or null: live electron-steps, distance
                            // evaluations, recombinations, steps simulated
    int32_t h_cap;          // LDS hole capacity (multiple of 4)
    int32_t e_cap;          // LDS electron capacity (multiple of 64, >= n_e0)
    int32_t g_cap;          // LDS cell-start capacity (>= G^3 + 1, multiple of 4)
    int32_t wave_search_max;  // queues up to this long are searched one electron per wave (<= 64)
    int32_t lookahead_max;    // with at most this many live electrons, quiet steps are skipped
    int32_t quiet_window;     // 1: the kernels built with the quiet window (after quiet
                              // steps, the rest of a Philox block tested at once)
    StateDump dump;
    // global-memory form (clusters beyond one workgroup's LDS): workgroup b keeps its
    // cluster's holes and cell starts (and, unless e_lds, its electron arrays) in
    // gscratch + b * gstride
    unsigned char* gscratch;
    int64_t gstride;
    int32_t e_lds;            // electron arrays in LDS after the exchange slots
    // 1: some run of the launch is isothermal with one localisation channel, and
    // the kernels built with the fixed-threshold path run it; 0: the kernels
    // without that path (same counts, shorter code for ramped or two-channel runs)
    int32_t fixed_thr;
};

// Event-mode launch: one cluster per workgroup, one event per step.
struct EventArgs {
    const RunDesc* runs;
    const int32_t* c_run;   // [n_clusters]
    const int32_t* c_sim;   // [n_clusters]
    int32_t n_clusters;
    uint32_t key0, key1;
    int32_t max_steps;      // rows per cluster
    uint8_t* lum;           // [n_clusters][max_steps]: 1 for a recombination step
    double* t_axis;         // [n_clusters][max_steps] time after the step, or null
    double* e_ratio;        // [n_clusters][max_steps] n_e / N_e after the step, or null
    int32_t* steps;         // [n_clusters] rows used; -1 - rows when max_steps ran out,
                            // kEventFillUnsupported when a filling event would continue
    uint64_t* work;         // [n_clusters][4] or null
    int32_t h_cap, e_cap, g_cap;
    int32_t wave_search_max;
    // global-memory form (clusters beyond LDS), as LaunchArgs: persistent over a
    // ticket, workgroup b's holes and cell starts (and, unless e_lds, its electron
    // arrays) at gscratch + b * gstride
    uint32_t* queue;
    unsigned char* gscratch;
    int64_t gstride;
    int32_t e_lds;
};
constexpr int32_t kEventFillUnsupported = -0x7fffffff;

// Lab experiments (TL_lab / ISO_lab): event mode with dose filling, recording
// n_e / N_e at observation times.  One cluster per workgroup.
struct LabArgs {
    const RunDesc* runs;
    const int32_t* c_run;
    const int32_t* c_sim;
    int32_t n_clusters;
    uint32_t key0, key1;
    const double* obs;          // observation times of every run, concatenated (ascending per run)
    const int32_t* obs_offset;  // [n_runs] first observation of each run
    const int32_t* obs_count;   // [n_runs]
    int32_t max_obs;            // ratio row stride
    double* ratio;              // [n_clusters][max_obs]
    int32_t* status;            // [n_clusters]: 0; -1 step budget spent; -2 capacity exceeded
    uint64_t* work;             // [n_clusters][4] live electron-steps, fills, recombinations, steps
    int64_t max_steps;
    int32_t h_cap, e_cap;
    uint32_t options;           // MCL_LAB_* bits
};
int lab_lds_bytes(int h_cap, int e_cap);

constexpr int kWaveHost = 64;   // wavefront width of gfx950

// LDS bytes the fixed-dt kernel needs for the given capacities.
inline int fixed_dt_lds_bytes(int h_cap, int e_cap, int g_cap)
{
    return 16 * h_cap + 4 * g_cap + 16 * e_cap;
}

// Global-memory form: scratch bytes per workgroup (a multiple of 256), and the
// LDS of the cluster arrays kept there (the per-wave extra is added at launch).
inline int64_t fixed_dt_global_stride(int h_cap, int e_cap, int g_cap, bool e_lds)
{
    const int64_t b = 16 * (int64_t)h_cap + 4 * (int64_t)g_cap + (e_lds ? 0 : 16 * (int64_t)e_cap);
    return (b + 255) / 256 * 256;
}
inline int fixed_dt_global_lds(int e_cap, bool e_lds) { return e_lds ? 16 * e_cap : 0; }

// Cells per axis of a cluster's hole grid: about holes_per_cell holes a cell.
inline int grid_dim(int n_h, double holes_per_cell)
{
    if (n_h <= 0) return 1;
    int G = 1;
    while ((double)(G + 1) * (G + 1) * (G + 1) * holes_per_cell <= (double)n_h && G < 31) ++G;
    return G;
}

// LDS beyond fixed_dt_lds_bytes for a workgroup of `waves` wavefronts: the
// per-wave exchange slots (2 x waves words) and the persistent ticket.
constexpr int fixed_dt_waves_lds(int waves)
{
    return waves > 1 ? ((4 * (2 * waves + 1) + 15) / 16) * 16 : 0;
}

// waves = wavefronts per cluster (1, 2, 4 or 8); lds_bytes = fixed_dt_lds_bytes of the
// launch (the per-wave extra is added here)
hipError_t launch_fixed_dt(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                           hipStream_t stream);
hipError_t launch_reset_tickets(uint32_t* tickets, int n, hipStream_t stream);
hipError_t prepare_fixed_dt(int lds_bytes, int max_lds);
hipError_t fixed_dt_occupancy(int lds_bytes, int waves, int* blocks_per_cu);
// The global-memory form: persistent launches only (a.queue set, grid at most the
// slices of a.gscratch); lds_bytes = fixed_dt_global_lds.
hipError_t launch_fixed_dt_global(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                  hipStream_t stream);
hipError_t prepare_fixed_dt_global(int max_lds);
hipError_t fixed_dt_global_occupancy(int lds_bytes, int waves, int* blocks_per_cu);
// The builds with the quiet window (mc_fixed_dt_window.hip,
// mc_fixed_dt_global_window.hip); the functions above call them when
// a.quiet_window is set and take both builds into the limits and occupancy.
hipError_t prepare_fixed_dt_window(int lds_bytes, int max_lds);
hipError_t launch_fixed_dt_window(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                  hipStream_t stream);
hipError_t fixed_dt_window_occupancy(int lds_bytes, int waves, int* blocks_per_cu);
hipError_t prepare_fixed_dt_global_window(int max_lds);
hipError_t launch_fixed_dt_global_window(const LaunchArgs& a, int grid, int lds_bytes, int waves,
                                         hipStream_t stream);
hipError_t fixed_dt_global_window_occupancy(int lds_bytes, int waves, int* blocks_per_cu);
hipError_t launch_event(const EventArgs& a, int lds_bytes, hipStream_t stream);
// The event kernel's global-memory form: grid workgroups over a.queue's ticket.
hipError_t launch_event_global(const EventArgs& a, int grid, int lds_bytes, hipStream_t stream);
hipError_t event_global_occupancy(int lds_bytes, int* blocks_per_cu);
hipError_t launch_lab(const LabArgs& a, int lds_bytes, hipStream_t stream);
hipError_t prepare_lab(int lds_bytes);
hipError_t prepare_event(int lds_bytes);

}  // namespace mcl
