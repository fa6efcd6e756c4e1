/*
 * mcl.h — C ABI of the MI355X trap/charge-kinetics Monte Carlo library (libmcl.so).
 *
 * The reference has no FFI: its hot path is the Python step loop in
 * src/class/simulate.py:46-92 driving TLTrapSim (src/class/tl_trap_lab.py:24-60),
 * Box (src/class/engine.py:83-188) and Physics (src/class/engine.py:44-77).
 * Driving that loop one step at a time from Python cannot run at GPU speed, so
 * the boundary sits one level up: one call advances every (run, sim) cluster
 * of a sweep through all of its steps.  The Python package mirrors the
 * reference's classes on top of these entry points (ctypes).
 *
 * Conventions
 *   - Every entry point returns MCL_OK (0) or a negative MCL_E* status; the
 *     message of the last failure on a context is mcl_last_error(ctx).
 *   - Plain pointers and sizes only.  Output buffers are caller-owned: host
 *     memory by default, device memory when MCL_F_DEVICE_PTRS is set.
 *   - A context owns one HIP stream and its device buffers; it is not
 *     thread-safe (use one context per host thread).
 *   - Clusters are ordered run-major then sim, the loop order of
 *     simulate.py:36-46 and the row order of its CSV (simulate.py:100-110).
 */
#ifndef MCL_H
#define MCL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCL_ABI_VERSION 1

/* status codes (negated errno values) */
#define MCL_OK         0
#define MCL_EDEVICE   (-5)   /* HIP runtime error                                  */
#define MCL_ENOMEM    (-12)  /* device allocation failed                           */
#define MCL_EINVAL    (-22)  /* invalid argument or parameter set                  */
#define MCL_ERANGE    (-34)  /* cluster too large for the kernel's LDS budget      */
#define MCL_EOVERFLOW (-75)  /* output buffer too small for the requested bins     */

/* flags for mcl_run_fixed_dt */
#define MCL_F_DEVICE_PTRS  0x1u  /* counts / live_steps are device pointers        */
#define MCL_F_PERSISTENT   0x2u  /* persistent kernel pulling clusters from a queue */
#define MCL_F_GRAPH        0x4u  /* capture the launch in a hipGraph and replay it  */

/* Philox4x32-10 stream tags: counter word 3 = (sim << 4) | tag */
#define MCL_TAG_GEOM_E 0u   /* electron coordinates, counter (0, e, run, sim|tag)   */
#define MCL_TAG_GEOM_H 1u   /* hole coordinates, counter (0, h, run, sim|tag)       */
#define MCL_TAG_FIRE   2u   /* recombination draw, counter (n>>2, e, run, sim|tag)  */
#define MCL_TAG_CHAN   3u   /* localisation-channel draw, same counter layout       */
#define MCL_TAG_RETRY0 4u   /* conflict retry after round r: tag RETRY0 + r         */

/* Conflict-resolution rounds per step: round 0 resolves the electrons that
 * fired; each later round retries the previous round's losers on their new
 * nearest hole.  Losers left after the last round wait for the next step. */
#define MCL_MAX_ROUNDS 12

/* Index sentinel for "no live hole" (n_h reached 0). */
#define MCL_NO_HOLE 0xFFFFu

/*
 * One parameter set ("run") with its sims = independent clusters.
 * Field meanings follow the reference YAML keys:
 *   physics_fp: alpha [1/m], b, s [1/s], E_cb, E_loc_1, E_loc_2 [eV], D0, Retrap,
 *               k_b [eV/K], D [Gy/s]                       (engine.py:48-60)
 *   exp_type_fp: rho_prime, holes, boundary_factor, N_e, e_ratio_start,
 *               T_start [degC], T_rate [K/s]               (tl_trap_lab.py:33-39,
 *                                                           simulate.py:40-44)
 * plus the fixed-dt schedule: dt [s] and n_steps.  Temperature at step n is
 * T_n = (T_start + 273.15) + T_rate * (n * dt), the start-of-step temperature
 * used by simulate.py:53-54.
 */
typedef struct mcl_run_params {
    double alpha, b, s, E_cb, E_loc_1, E_loc_2, D0, Retrap, k_b, D;
    double rho_prime, holes, boundary_factor, N_e, e_ratio_start;
    double T_start, T_rate, dt;
    int32_t n_steps;
    int32_t sims;
} mcl_run_params;

/* Geometry derived from one run (TLTrapSim.__init__, tl_trap_lab.py:33-39;
 * Box.seed, engine.py:124-129). */
typedef struct mcl_geometry {
    double rho;        /* hole density [1/m^3] = rho_prime * 3/(4 pi) * alpha^3      */
    double side_m;     /* core cube side d = (holes/rho)^(1/3) [m]                   */
    int32_t n_e0;      /* int(N_e * e_ratio_start) electrons in [0, d)^3             */
    int32_t n_h;       /* int(holes * bf^3) holes in [0, bf d)^3                     */
} mcl_geometry;

typedef struct mcl_stats {
    double   kernel_ms;            /* device time of the MC launch (HIP events)      */
    uint64_t live_electron_steps;  /* sum over clusters and steps of live electrons  */
    uint64_t recombinations;       /* total recombination events                     */
    uint64_t clusters;             /* clusters simulated                             */
    uint64_t cluster_steps;        /* sum over clusters of simulated steps           */
} mcl_stats;

/* ---- context ------------------------------------------------------------- */
int         mcl_abi_version(void);
int         mcl_create(void** ctx_out, int hip_device);
/* A context keeps the device buffers of finished runs and destroyed plans (up to
 * 512 MiB) for later ones; mcl_destroy frees them. */
void        mcl_destroy(void* ctx);
const char* mcl_last_error(const void* ctx);
/* Launch subsequent work on an existing HIP stream (e.g. torch's current
 * stream); NULL restores the context's own stream. */
int         mcl_set_stream(void* ctx, void* hip_stream);
/* Largest cluster (electrons, holes) one workgroup can hold in LDS.  Fixed-dt
 * runs (mcl_run_fixed_dt, plans, mcl_run_cluster, mcl_cluster_state) and
 * event-mode runs take larger clusters, up to 32768 (fixed-dt) or 65535 (event)
 * electrons and 65534 holes, in a global-memory form: holes and cell starts in a
 * scratch slice per resident workgroup (16 B per hole slot, 4 B per cell),
 * electrons in LDS while 16 B each fit; same results, lower speed.  Lab runs and
 * mcl_run_fixed_dt_noalloc stay within this limit. */
int         mcl_cluster_limits(const void* ctx, int32_t* max_electrons, int32_t* max_holes);
/* Kernel tuning knobs of a context, applied to plans made afterwards.  Keys:
 *   "holes_per_cell"  mean holes per cell of each cluster's nearest-hole grid
 *                     (default 2; results do not depend on it, only speed).
 *   "persist_frac"    persistent workgroups per size class, as a fraction of
 *                     what fits on the device at once (default 1).
 *   "wave_search_max" electrons needing a new nearest hole in one step are
 *                     searched by the whole wave, up to four at a time, up
 *                     to this many, lane-parallel beyond (default 32, at
 *                     most 64).
 *   "lookahead_max"   with at most this many live electrons in a cluster, the
 *                     wave tests the next 64 steps at once and skips those in
 *                     which no electron can fire (default 8; 0 turns it off).
 *   "waves_per_cluster" wavefronts per fixed-dt cluster: 1, 2, 4 or 8 (the waves
 *                     split the cluster's 64-electron rows, at most 64 rows a
 *                     wave, so clusters above 4096 electrons take at least the
 *                     waves that cover them), or 0 (default): each launch, or
 *                     each size class of a persistent launch, takes the count
 *                     that keeps the most waves resident.
 *   "global_memory"   1: fixed-dt plans and event-mode runs keep every
 *                     cluster's holes in the global-memory form (results
 *                     identical, slower); 0 (default): only clusters that do
 *                     not fit LDS.
 *   "fixed_thresholds" 1 (default): a fixed-dt launch (or size class) with an
 *                     isothermal one-localisation-channel run takes the kernels
 *                     that keep each electron's fire threshold in its w slot;
 *                     0: the kernels that compute it from w every step (results
 *                     identical; ramped and two-channel launches always take
 *                     those).
 *   "quiet_window"    1 (default): fixed-dt launches (or size classes) with a
 *                     ramped run take the kernels with the quiet window: after
 *                     two steps in which no electron fired, the steps left in
 *                     the Philox block (up to three) are tested at once and
 *                     skipped together when no electron fires in them; 0: the
 *                     kernels without it (results identical).
 * No counterpart in the reference (it has no kernels); the autotuner
 * (mcluminescence_amd/autotune.py) sets them. */
int         mcl_tune_set(void* ctx, const char* key, double value);
int         mcl_tune_get(const void* ctx, const char* key, double* value);

/* ---- geometry (host only) ------------------------------------------------ */
/* Replaces the geometry part of TLTrapSim.__init__ (tl_trap_lab.py:33-39). */
int mcl_run_geometry(const mcl_run_params* run, mcl_geometry* out);

/* ---- fixed-dt Monte Carlo -------------------------------------------------- */
/*
 * Replaces the run x sim x step loop of simulate.py:36-92 (one event per step)
 * by the fixed-dt multi-recombination step (per live electron one Bernoulli
 * draw with p = 1 - exp(-k dt), k = s e^{-E_cb/kT} + b e^{-E_loc/kT - alpha r};
 * conflicts over a shared nearest hole go to the lowest electron index, the
 * rule of src/est_params/functions.py:228-239).
 *
 *   counts     [total_clusters][row_len] recombinations per bin; bin of step n is
 *              n * n_bins / n_steps (n_bins <= n_steps of every run) and
 *              row_len = n_bins.  n_bins = 0 gives per-step counts (the
 *              reference's Lum column) with row_len = the longest n_steps, each
 *              row zero-padded after its own n_steps.
 *   live_steps [total_clusters] live electron-steps per cluster, or NULL.
 * The Dose filling path (D > 0) is not part of the fixed-dt kernel: runs with
 * D != 0 are rejected with MCL_EINVAL, matching simulate.py:42 which always
 * runs with dose_rate 0.
 */
int mcl_run_fixed_dt(void* ctx, const mcl_run_params* runs, int32_t n_runs,
                     uint64_t seed, int32_t n_bins, uint32_t flags,
                     uint32_t* counts, uint64_t* live_steps, mcl_stats* stats);

/*
 * Plans: upload the parameter sets once, launch many times.  A plan owns the
 * device-side run descriptors and cluster map (and, with MCL_F_GRAPH, a
 * captured hipGraph).  mcl_plan_launch is asynchronous on the context's
 * stream and takes DEVICE pointers for counts [total_clusters][row_len] and
 * live_steps [total_clusters] (may be NULL); mcl_sync waits for it.
 */
typedef struct mcl_plan_info {
    int64_t total_clusters;
    int32_t n_bins;
    int32_t row_len;          /* counts row stride                               */
    int32_t lds_bytes;        /* dynamic LDS per workgroup                       */
    int32_t max_electrons;    /* largest n_e0 in the plan                        */
    int32_t max_holes;        /* largest n_h in the plan                         */
    int32_t grid;             /* workgroups launched                             */
    int32_t block;            /* threads per workgroup                           */
} mcl_plan_info;

int  mcl_plan_fixed_dt(void* ctx, const mcl_run_params* runs, int32_t n_runs,
                       uint64_t seed, int32_t n_bins, uint32_t flags, void** plan_out);
int  mcl_plan_info_get(const void* plan, mcl_plan_info* info);
int  mcl_plan_launch(void* plan, uint32_t* counts_dev, uint64_t* live_steps_dev);

/* Work done by the last launch of a plan, summed over its clusters (waits for
 * the launch): the algorithmic units behind throughput and roofline figures. */
typedef struct mcl_work {
    uint64_t live_electron_steps;  /* live electrons summed over every simulated step   */
    uint64_t distance_evals;       /* electron-hole squared distances evaluated          */
    uint64_t recombinations;       /* electron-hole pairs removed                        */
    uint64_t cluster_steps;        /* steps simulated until each cluster emptied         */
    double   kernel_ms;            /* device time of the launch (HIP events)             */
} mcl_work;

int  mcl_plan_work(void* plan, mcl_work* out);
/* Device time of the last launch of this plan (HIP events on the launch stream). */
int  mcl_plan_last_ms(void* plan, double* ms);
void mcl_plan_destroy(void* plan);
int  mcl_sync(void* ctx);

/*
 * Event-mode Monte Carlo: the reference's loop itself (simulate.py:46-92), one
 * event per step, for its exact step / lum / electron_ratio rows.  Per step the
 * total rate K = sum of every live electron's k_e (engine.py:65-77, channel as in
 * the fixed-dt step) gives dt_rec = -ln(U) / K (the earliest of the electrons'
 * exponential waits, tl_trap_lab.py:48-51), dt = min(dt_fill, dt_rec, dt_cap) with
 * dt_cap = dt / T_rate (the run's dt field holds max_dt; no cap at T_rate = 0),
 * and a recombination picks electron e with probability k_e / K; it and its
 * nearest hole are removed (Box.remove_pair, engine.py:154-175).
 *   duration  [n_runs] seconds: the loop runs while t <= duration and stops after
 *             the step that reaches it (simulate.py:51,91).
 *   lum       [total_clusters][max_steps] 1 for a recombination step, else 0
 *   t_axis    [total_clusters][max_steps] time after each step (x_ax), or NULL
 *   e_ratio   [total_clusters][max_steps] n_e / N_e after each step, or NULL
 *   steps     [total_clusters] rows used (negative: -1 - rows when max_steps ran
 *             out); rows after the last are zero, as in the reference's padding.
 * Returns MCL_EOVERFLOW when a cluster needs more than max_steps steps (the
 * reference raises IndexError), MCL_EINVAL for D != 0 or a filling event that
 * would continue the run.  The run's n_steps must be >= 1 and is not used.
 */
int mcl_run_event(void* ctx, const mcl_run_params* runs, int32_t n_runs, const double* duration,
                  uint64_t seed, int32_t max_steps, uint32_t flags, uint8_t* lum,
                  double* t_axis, double* e_ratio, int32_t* steps, mcl_stats* stats);

/*
 * Lab experiments with dose filling: the loops of TLTrapSim.TL_lab and ISO_lab
 * (tl_trap_lab.py:65-179).  Each run is one experiment: T_start [degC] and T_rate
 * [K/s] give T = (T_start + 273.15) + T_rate t, D is the dose rate and
 * e_ratio_start the initial fill; dt and n_steps are not used.  Events as in
 * mcl_run_event without a step cap, plus filling at lambda = (D/D0)(N_e - n_e)
 * (tl_trap_lab.py:53-60): one electron-hole pair per fill (Box.add_electron,
 * engine.py:133-152).  Without MCL_LAB_EXACT_FILL the new electron's nearest hole
 * is searched among the older holes and no other electron looks at the new hole,
 * the reference's update; with it every nearest hole is exact after a fill.
 * The trapped fraction n_e / N_e after the first step that reaches each
 * observation time is written to ratio.
 *   obs_times  all runs' observation times concatenated, ascending within a run
 *   obs_count  [n_runs] observations per run (>= 1)
 *   options    MCL_LAB_* bits below (0: the class-based loops of tl_trap_lab.py)
 *   ratio      [total_clusters][max over runs of obs_count]
 *   status     [total_clusters] 0, -1 (max_steps spent), -2 (capacity), or NULL
 * Returns MCL_EOVERFLOW when a cluster needs more than max_steps steps.
 *
 * The other bits reproduce the legacy lab loops sim_lab_TL_residuals /
 * sim_lab_TL_residuals_iso (src/est_params/functions.py:270-434):
 *   MCL_LAB_SHARED_CHANNEL   one localisation-channel draw per step shared by all
 *                            electrons: E_loc_2 iff U <= Retrap (functions.py:125-128)
 *   MCL_LAB_RETRAP           after each recombination, with probability Retrap, one
 *                            new electron-hole pair (functions.py:328-331; the TL loop)
 *   MCL_LAB_FIRST_ELECTRON   the recombining electron is the oldest one (creation
 *                            order) instead of one drawn with probability k_e / K:
 *                            recomber as those loops call it (functions.py:241-261,
 *                            its np.where over a (1, n_e) mask yields index 0)
 *   MCL_LAB_LEGACY_GEOMETRY  holes = int(holes) + int(int(holes)/d^3 * ((bf d)^3 - d^3))
 *                            with d = (int(holes)/rho)^(1/3) (initialize_box_bg,
 *                            functions.py:51-80)
 *   MCL_LAB_LEGACY           all of these plus MCL_LAB_EXACT_FILL (the legacy fill
 *                            recomputes every nearest hole, functions.py:150-161)
 */
#define MCL_LAB_EXACT_FILL       0x1u
#define MCL_LAB_SHARED_CHANNEL   0x2u
#define MCL_LAB_RETRAP           0x4u
#define MCL_LAB_FIRST_ELECTRON   0x8u
#define MCL_LAB_LEGACY_GEOMETRY  0x10u
#define MCL_LAB_LEGACY           0x1Fu
int mcl_run_lab(void* ctx, const mcl_run_params* runs, int32_t n_runs, const double* obs_times,
                const int32_t* obs_count, uint32_t options, uint64_t seed, int64_t max_steps,
                double* ratio, int32_t* status, mcl_stats* stats);

/*
 * Glow curves, one per run: the binning and smoothing of plots.py:39-48
 * (hist_and_smooth) on the data path of plot_TL (plots.py:50-66), computed from
 * the per-step counts of a sweep.  For run r with S = sims and first cluster c0:
 *   t_i    t_axis[c0][i] (the first sim's time axis, plots.py:59), or with t_axis
 *          NULL the fixed-dt axis (i + 1) dt for i < n_steps and 0 beyond
 *   w_i    (sum of count[c][i] over the run's clusters) / S   (plots.py:60)
 *   T_i    T_start + T_rate t_i for every step with t_i > 0  (plots.py:61-63)
 *   hist   weighted histogram of T_i over the edges j * bin_width, j = 0 ..
 *          ceil((max T + bin_width) / bin_width) - 1 (np.arange), bins
 *          [e_j, e_j+1) with the last one closed, divided by bin_width
 *   curve  boxcar mean over k = max(1, int(win / bin_width)) bins, 'valid' mode
 *          (np.convolve with ones(k) / k, plots.py:20-22; when k exceeds the bin
 *          count every point is the total over k)
 *   counts     [total_clusters][row_stride]: u32 (count_bytes 4, fixed-dt counts)
 *              or u8 (count_bytes 1, event-mode lum)
 *   t_axis     [total_clusters][row_stride] f64, or NULL
 *   curve      [n_runs][curve_stride] f64; curve_len [n_runs] points written (0
 *              where no step has t > 0 or every T is <= 0: the reference raises)
 * With MCL_F_DEVICE_PTRS counts, t_axis, curve and curve_len are device
 * pointers; the call waits for the kernel either way.  A histogram with more
 * than MCL_GLOW_MAX_BINS bins fails with MCL_ERANGE, a curve longer than
 * curve_stride with MCL_EOVERFLOW (curve_len then holds the lengths needed).
 * Sums over sims are exact (64-bit integers); each point is one division.
 */
#define MCL_GLOW_MAX_BINS 8192
int mcl_glow_curve(void* ctx, const mcl_run_params* runs, int32_t n_runs, const void* counts,
                   int32_t count_bytes, int64_t row_stride, const double* t_axis,
                   double bin_width, double win, uint32_t flags, double* curve,
                   int64_t curve_stride, int32_t* curve_len, mcl_stats* stats);

/*
 * Allocation-free variant of mcl_run_fixed_dt, for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).  It has
 * no counterpart in the reference.  Same runs, seed and n_bins as
 * mcl_run_fixed_dt with MCL_F_DEVICE_PTRS (one workgroup per cluster, no
 * persistent or graph launch) and the same counts and live electron-steps, but
 * no allocation of any kind: the run descriptors, cluster map and work counters
 * go to a static device array of MCL_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends), staged from a fixed host block.
 *   counts_dev      DEVICE pointer [total_clusters][row_len] (caller-owned)
 *   live_steps_dev  DEVICE pointer [total_clusters], or NULL
 * A call that needs more scratch than that (192 B per run plus 40 B per
 * cluster, each part rounded up to 256 B) fails with MCL_ERANGE before any
 * device work.  Every call on one GPU uses that GPU's single scratch array, so
 * calls that overlap on one GPU (several host threads or contexts) race; calls
 * are synchronous on the context's stream, so calls from one thread never
 * overlap.  Nothing in the scratch needs clearing: every call writes all of it
 * that its kernel reads.
 */
#define MCL_NOALLOC_SCRATCH_BYTES (4u << 20)
int mcl_run_fixed_dt_noalloc(void* ctx, const mcl_run_params* runs, int32_t n_runs,
                             uint64_t seed, int32_t n_bins, uint32_t* counts_dev,
                             uint64_t* live_steps_dev, mcl_stats* stats);

/*
 * Counts of one cluster of a sweep: cluster (run_index, sim) of the sweep that
 * holds `run` at position run_index, identical to its row in mcl_run_fixed_dt,
 * without simulating the sweep's other clusters.  For the one-cluster object
 * TLTrapSim (tl_trap_lab.py:24-43), one simulation of one run.  Host pointers:
 * counts [row_len] (n_bins, or run->n_steps for n_bins = 0), live_steps [1] or NULL.
 */
int mcl_run_cluster(void* ctx, const mcl_run_params* run, int32_t run_index, int32_t sim,
                    uint64_t seed, int32_t n_bins, uint32_t* counts, uint64_t* live_steps);

/*
 * Per-set curves from per-step counts (device pointers, asynchronous on the
 * context's stream): for set p and bin j,
 *   out[p][j] = sum over its sims k < sims and steps n in [edges[j], edges[j+1])
 *               of counts[p * sims + k][n]
 * with counts rows of row_stride u32 (mcl_run_fixed_dt / a plan with n_bins = 0)
 * and edges a host array of n_bins + 1 increasing step indices from >= 0, the
 * last <= row_stride and <= 16384.  The surrogate's isothermal curves use
 * log-spaced edges.  No counterpart in the reference (its binning is by
 * temperature, plots.py:39-47; mcl_glow_curve).  MCL_ERANGE beyond MCL_BIN_MAX
 * bins or 16384 steps.
 */
#define MCL_BIN_MAX 512
int mcl_bin_counts(void* ctx, const uint32_t* counts_dev, int64_t n_sets, int32_t sims,
                   int64_t row_stride, const int32_t* edges, int32_t n_bins, uint64_t* out_dev);

/*
 * State of cluster (run_index, sim) as the kernel holds it after max_steps
 * steps (0 = the initial state): coordinates [nm] in float32, which electrons
 * and holes are left, the incrementally maintained nearest live hole per
 * electron (argmin, first index on ties) and w = exp(-alpha r).  Replaces
 * Box.seed, Box._rebuild and Box.remove_pair's bookkeeping (engine.py:113-175)
 * for inspection.  Host pointers: e_xyz [n_e0][3], h_xyz [n_h][3] (a removed
 * hole's x reads 3e38), e_live [n_e0], h_alive [n_h], nn [n_e0], w [n_e0].
 */
int mcl_cluster_state(void* ctx, const mcl_run_params* run, int32_t run_index,
                      int32_t sim, uint64_t seed, int32_t max_steps,
                      float* e_xyz, float* h_xyz, uint8_t* e_live, uint8_t* h_alive,
                      uint32_t* nn, float* w);

#ifdef __cplusplus
}
#endif
#endif /* MCL_H */
