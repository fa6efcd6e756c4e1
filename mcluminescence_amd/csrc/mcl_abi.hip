// mcl_abi.hip — the C ABI of libmcl.so (declared in include/mcl.h).
//
// Host side of the drop-in boundary: validates parameter sets the way the
// reference's classes would fail on them, derives each run's geometry
// (tl_trap_lab.py:33-39, engine.py:124-129), uploads descriptors once per plan
// and launches the HIP kernels on the context's stream.  No CPU fallback: if
// no HIP device is usable every entry point fails with MCL_EDEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mcl.h"
#include "mcl_ctx.h"
#include "mcl_internal.h"

// Device memory of mcl_run_fixed_dt_noalloc (run descriptors, cluster map and
// per-cluster work counters of one call): a file-scope array, one copy per GPU,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, once per device in mcl_create, and only that address is
// used (kernel arguments, copies).
static __device__ __align__(256) unsigned char scratch[MCL_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices] = {};   // device address per HIP device
hipError_t g_scratch_err[kMaxScratchDevices] = {};

// With `device` current: record the device address of its scratch copy.
void init_scratch(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_dev[device] = (unsigned char*)ptr;
    else (void)hipGetLastError();        // reported by the entry point that needs it
}

unsigned char* scratch_for(int device, hipError_t* err)
{
    if (device < 0 || device >= kMaxScratchDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    *err = g_scratch_err[device];
    return g_scratch_dev[device];
}
}  // namespace

namespace mcl {

namespace {
constexpr size_t kPoolCap = size_t(512) << 20;   // idle bytes a context keeps at most
}

hipError_t pool_alloc(Ctx* c, void** p, size_t bytes)
{
    bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
    {
        std::lock_guard<std::mutex> lock(c->pool_mu);
        // the smallest idle block that fits and wastes at most about half of itself
        size_t best = c->pool_free.size();
        for (size_t i = 0; i < c->pool_free.size(); ++i) {
            const size_t b = c->pool_free[i].first;
            if (b >= bytes && b <= 2 * bytes + 65536 &&
                (best == c->pool_free.size() || b < c->pool_free[best].first))
                best = i;
        }
        if (best < c->pool_free.size()) {
            const auto blk = c->pool_free[best];
            c->pool_free[best] = c->pool_free.back();
            c->pool_free.pop_back();
            c->pool_cached -= blk.first;
            c->pool_live[blk.second] = blk.first;
            *p = blk.second;
            return hipSuccess;
        }
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(c->pool_mu);
    c->pool_live[*p] = bytes;
    return hipSuccess;
}

void pool_release(Ctx* c, void* p, bool idle)
{
    if (!p) return;
    if (!idle) (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lock(c->pool_mu);
    const auto it = c->pool_live.find(p);
    if (it == c->pool_live.end()) return;
    const size_t bytes = it->second;
    c->pool_live.erase(it);
    if (c->pool_cached + bytes > kPoolCap) {
        (void)hipFree(p);
        return;
    }
    c->pool_free.emplace_back(bytes, p);
    c->pool_cached += bytes;
}

void pool_clear(Ctx* c)
{
    std::lock_guard<std::mutex> lock(c->pool_mu);
    for (const auto& blk : c->pool_free) (void)hipFree(blk.second);
    c->pool_free.clear();
    c->pool_cached = 0;
}

}  // namespace mcl

namespace {

using mcl::Ctx;
using mcl::fail;


struct Plan {
    Ctx* ctx = nullptr;
    mcl::RunDesc* d_runs = nullptr;
    int32_t* d_crun = nullptr;
    int32_t* d_csim = nullptr;
    uint64_t* d_work = nullptr;   // [total][4] per-cluster work counters
    uint32_t key0 = 0, key1 = 0;
    int64_t total = 0;
    int32_t n_bins = 0;
    int32_t row_len = 0;          // counts row stride
    int32_t h_cap = 0, e_cap = 0, g_cap = 0, lds = 0;
    int32_t max_e = 0, max_h = 0;
    int32_t waves = 1;            // wavefronts per cluster of the one-workgroup-per-cluster launch
    int32_t fixed_thr = 0;        // some run keeps its fire thresholds (keeps_thresholds)
    int32_t window = 0;           // some run takes the quiet window (uses_window)
    uint32_t flags = 0;
    // clusters beyond one workgroup's LDS (global-memory form): the plain launch is
    // persistent, ggrid workgroups with a scratch slice of gstride bytes each
    bool global = false;
    int32_t e_lds = 0;
    int64_t gstride = 0;
    int32_t ggrid = 0;
    size_t class_gbytes = 0;               // slices of the global-form classes
    unsigned char* d_gscratch = nullptr;   // slices of the plain launch or of the classes
    uint32_t* d_gticket = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    // persistent launch: one class kernel per cluster-size class
    struct Class {
        int32_t first = 0, count = 0;   // slice of d_order (clusters, most work first)
        int32_t h_cap = 0, e_cap = 0, g_cap = 0, lds = 0, grid = 0;
        int32_t max_e = 0, max_h = 0;
        int32_t waves = 1;
        int32_t fixed_thr = 0;          // some cluster's run keeps its fire thresholds
        int32_t window = 0;             // some cluster's run takes the quiet window
        bool global = false;            // global-memory form: scratch at goff, gstride a slice
        int32_t e_lds = 0;
        int64_t gstride = 0;
        size_t goff = 0;
    };
    std::vector<Class> classes;
    int32_t* d_order = nullptr;
    uint32_t* d_tickets = nullptr;
    std::vector<hipStream_t> side;      // fork streams for concurrent class kernels
    std::vector<hipEvent_t> join;       // one per class
    hipEvent_t fork = nullptr;
    // graph of the persistent launch, instantiated for one (counts, live) pair
    hipGraphExec_t gexec = nullptr;
    uint32_t* g_counts = nullptr;
    uint64_t* g_live = nullptr;
};

constexpr double kPi = 0x1.921fb54442d18p+1;

void geometry(const mcl_run_params& p, mcl_geometry& g)
{
    // tl_trap_lab.py:33-34 and engine.py:127 (int() truncation)
    g.rho = p.rho_prime * (3.0 / (4.0 * kPi) * std::pow(p.alpha, 3.0));
    g.side_m = std::pow(p.holes / g.rho, 1.0 / 3.0);
    g.n_e0 = (int32_t)(p.N_e * p.e_ratio_start);
    g.n_h = (int32_t)(p.holes * std::pow(p.boundary_factor, 3.0));
}

// The legacy stack's box (initialize_box_bg, functions.py:51-80): int(holes) core
// holes set the side, and the boundary shell adds int(density * shell volume).
void legacy_geometry(const mcl_run_params& p, mcl_geometry& g)
{
    g.rho = p.rho_prime * (3.0 / (4.0 * kPi) * std::pow(p.alpha, 3.0));
    const double h = std::trunc(p.holes);
    const double d = std::pow(h / g.rho, 1.0 / 3.0);
    const double db = d * p.boundary_factor;
    const double vol = d * d * d;
    const double shell = db * db * db - vol;
    g.side_m = d;
    g.n_e0 = (int32_t)(p.N_e * p.e_ratio_start);
    g.n_h = (int32_t)h + (int32_t)(h / vol * shell);
}

bool finite_all(const mcl_run_params& p)
{
    const double v[] = {p.alpha, p.b, p.s, p.E_cb, p.E_loc_1, p.E_loc_2, p.D0, p.Retrap, p.k_b, p.D,
                        p.rho_prime, p.holes, p.boundary_factor, p.N_e, p.e_ratio_start,
                        p.T_start, p.T_rate, p.dt};
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int validate_run(Ctx* c, const mcl_run_params& p, int r, int32_t n_bins, mcl_geometry& g)
{
    if (!finite_all(p)) return fail(c, MCL_EINVAL, "run %d: non-finite parameter", r);
    if (!(p.dt > 0.0)) return fail(c, MCL_EINVAL, "run %d: dt must be > 0 (got %g)", r, p.dt);
    if (p.n_steps <= 0) return fail(c, MCL_EINVAL, "run %d: n_steps must be > 0", r);
    if (p.sims < 0) return fail(c, MCL_EINVAL, "run %d: sims must be >= 0", r);
    if (p.D != 0.0)
        return fail(c, MCL_EINVAL,
                    "run %d: dose filling (D=%g) is not part of the fixed-dt kernel", r, p.D);
    if (!(p.k_b > 0.0)) return fail(c, MCL_EINVAL, "run %d: k_b must be > 0", r);
    if (!(p.alpha > 0.0) || !(p.rho_prime > 0.0) || !(p.holes > 0.0) || !(p.boundary_factor > 0.0))
        return fail(c, MCL_EINVAL, "run %d: alpha, rho_prime, holes, boundary_factor must be > 0", r);
    if (p.s < 0.0 || p.b < 0.0) return fail(c, MCL_EINVAL, "run %d: s and b must be >= 0", r);
    const double T0 = p.T_start + 273.15;
    const double T1 = T0 + p.T_rate * ((double)(p.n_steps - 1) * p.dt);
    if (!(T0 > 0.0) || !(T1 > 0.0))
        return fail(c, MCL_EINVAL, "run %d: temperature must stay > 0 K (%g .. %g K)", r, T0, T1);
    if (n_bins < 0) return fail(c, MCL_EINVAL, "run %d: n_bins must be >= 0", r);
    if (n_bins > p.n_steps)
        return fail(c, MCL_EOVERFLOW, "run %d: n_bins=%d must be in [0, n_steps=%d]", r, n_bins,
                    p.n_steps);
    geometry(p, g);
    if (g.n_e0 < 0 || g.n_h < 0) return fail(c, MCL_EINVAL, "run %d: negative particle count", r);
    if (g.n_e0 > 65535 || g.n_h > 65534)
        return fail(c, MCL_ERANGE, "run %d: %d electrons / %d holes exceed 16-bit indices", r,
                    g.n_e0, g.n_h);
    return MCL_OK;
}

// Runs whose clusters keep their fire thresholds in the w slots (isothermal, one
// localisation channel): a launch with any of them takes the kernels built with
// that path.
bool keeps_thresholds(const Ctx* c, const mcl::RunDesc& d)
{
    return c->fixed_thresholds != 0.0 && d.T_rate == 0.0 && !d.two_channel;
}

// Runs on a temperature ramp: a launch with one of them takes the kernels with the
// quiet window (long stretches in which no electron fires, before and after the
// glow peak; an isothermal decay fires in nearly every step until its last few
// electrons, which the look-ahead covers).
bool uses_window(const Ctx* c, const mcl::RunDesc& d)
{
    return c->quiet_window != 0.0 && d.T_rate != 0.0;
}

mcl::RunDesc make_desc(const mcl_run_params& p, const mcl_geometry& g, int run_id,
                       double holes_per_cell)
{
    mcl::RunDesc d{};
    d.side_nm = g.side_m * 1e9;
    d.side_b_nm = (g.side_m * p.boundary_factor) * 1e9;
    d.alpha_nm = p.alpha * 1e-9;
    d.s = p.s;
    d.b = p.b;
    d.E_cb = p.E_cb;
    d.E_loc_1 = p.E_loc_1;
    d.E_loc_2 = p.E_loc_2;
    d.k_b = p.k_b;
    d.T0_K = p.T_start + 273.15;
    d.T_rate = p.T_rate;
    d.dt = p.dt;
    d.two_channel = (p.E_loc_1 != p.E_loc_2) && (p.Retrap > 0.0) && (p.Retrap < 1.0);
    d.always_chan2 = p.Retrap >= 1.0;
    d.ret_thr = d.two_channel ? (uint64_t)(p.Retrap * 0x1p32) : 0u;
    d.n_e0 = g.n_e0;
    d.n_h = g.n_h;
    d.n_steps = p.n_steps;
    d.run_id = run_id;
    d.grid = mcl::grid_dim(g.n_h, holes_per_cell);
    d.cell_nm = (float)(d.side_b_nm / d.grid);
    d.inv_cell = (float)(d.grid / d.side_b_nm);
    d.T_start_C = p.T_start;
    d.N_e = p.N_e;
    d.D = p.D;
    d.D0 = p.D0;
    d.duration = 0.0;
    d.retrap_thr = p.Retrap >= 1.0 ? (1ull << 32)
                                   : (p.Retrap > 0.0 ? (uint64_t)(p.Retrap * 0x1p32) : 0u);
    return d;
}

// idle: no kernel or copy of the plan is still running (its stream was synchronised).
void plan_free(Plan* p, bool idle)
{
    if (!p) return;
    Ctx* c = p->ctx;
    if (!idle) (void)hipDeviceSynchronize();
    mcl::pool_release(c, p->d_runs, true);
    mcl::pool_release(c, p->d_crun, true);
    mcl::pool_release(c, p->d_csim, true);
    mcl::pool_release(c, p->d_work, true);
    mcl::pool_release(c, p->d_order, true);
    mcl::pool_release(c, p->d_tickets, true);
    mcl::pool_release(c, p->d_gscratch, true);
    mcl::pool_release(c, p->d_gticket, true);
    if (p->gexec) hipGraphExecDestroy(p->gexec);
    for (hipStream_t s : p->side) hipStreamDestroy(s);
    for (hipEvent_t e : p->join) hipEventDestroy(e);
    if (p->fork) hipEventDestroy(p->fork);
    if (p->ev0) hipEventDestroy(p->ev0);
    if (p->ev1) hipEventDestroy(p->ev1);
    delete p;
}

// Wavefronts per cluster for `count` clusters of at most e_cap electrons that
// need lds bytes each: the tuning knob when set, otherwise the one of 1, 2, 4, 8
// (at most max_waves) that keeps the most waves resident (clusters resident at
// once, from the occupancy of that kernel, times waves), the fewest on a tie.  A
// wave gets at least one 64-slot row at the start (with 8-15 rows eight waves
// measured 2-20 % faster than four on launches of 16-64 clusters; with 5, four are
// 20 % faster than two).  Size classes that run side by side
// take at most four, so that a class of large clusters leaves wave slots on its
// CUs to the others (eight measured 20 % slower on an uneven sweep).  A wave
// keeps per-row flags (candidates, stale electrons, losers) in 64-bit masks, so it
// takes at most 64 rows: clusters above 4096 electrons get at least the waves
// that cover their rows (the knob included), above 32768 none do (MCL_ERANGE).
// per_cu: that kernel's workgroups per CU (global: of the global-memory kernels).
int choose_waves(Ctx* c, int lds, int e_cap, int64_t count, int max_waves, int* waves,
                 int* per_cu, bool global)
{
    const int rows = e_cap / 64;                            // 64-slot rows
    int need = 1;
    while (need < 8 && (rows + need - 1) / need > 64) need *= 2;
    if ((rows + need - 1) / need > 64)
        return fail(c, MCL_ERANGE, "clusters of %d electrons: a cluster holds at most %d (64 rows "
                    "of 64 slots for each of 8 waves)", e_cap, 8 * 64 * 64);
    const int forced = c->waves_per_cluster != 0 ? std::max((int)c->waves_per_cluster, need) : 0;
    int best = need, best_pc = 0;
    int64_t best_res = -1;
    for (int w : {1, 2, 4, 8}) {
        if (w < need) continue;
        if (forced != 0 && w != forced) continue;
        if (forced == 0 && w > need && (w > max_waves || w > rows)) break;
        if (lds + mcl::fixed_dt_waves_lds(w) > c->max_lds) continue;
        int pc = 0;
        const hipError_t e = global ? mcl::fixed_dt_global_occupancy(lds, w, &pc)
                                    : mcl::fixed_dt_occupancy(lds, w, &pc);
        if (e != hipSuccess) return fail(c, MCL_EDEVICE, "occupancy query: %s", hipGetErrorString(e));
        const int64_t res = (int64_t)w * std::min<int64_t>(count, (int64_t)std::max(pc, 1) * c->n_cu);
        if (res > best_res) {
            best = w;
            best_pc = pc;
            best_res = res;
        }
    }
    if (best_res < 0)
        return fail(c, MCL_ERANGE, "no workgroup of %d or more waves fits %d B of LDS per cluster",
                    need, lds);
    *waves = best;
    *per_cu = best_pc;
    return MCL_OK;
}

// Where the clusters of given capacities live: wholly in LDS when
// fixed_dt_lds_bytes and the largest exchange block fit a workgroup (and the
// global_memory knob is off); otherwise
// the global-memory form, with the electron arrays in LDS when they fit beside
// the largest exchange block.
// lds: the launch's dynamic LDS before the per-wave extra.
struct Layout {
    bool global = false;
    int32_t e_lds = 0;
    int32_t lds = 0;
    int64_t gstride = 0;
};

Layout layout_for(const Ctx* c, int h_cap, int e_cap, int g_cap)
{
    Layout L;
    L.lds = mcl::fixed_dt_lds_bytes(h_cap, e_cap, g_cap);
    if (L.lds + mcl::fixed_dt_waves_lds(8) <= c->max_lds && c->global_memory == 0.0) return L;
    L.global = true;
    L.e_lds = mcl::fixed_dt_global_lds(e_cap, true) + mcl::fixed_dt_waves_lds(8) <= c->max_lds;
    L.lds = mcl::fixed_dt_global_lds(e_cap, L.e_lds != 0);
    L.gstride = mcl::fixed_dt_global_stride(h_cap, e_cap, g_cap, L.e_lds != 0);
    return L;
}

// Persistent plans: bucket the clusters by electron count (powers of two from
// 128) so each class kernel sizes its LDS for its own clusters, order each class
// by expected work (electrons x steps, largest first) and give it as many
// persistent workgroups as fit on the device at once.
int build_classes(Ctx* c, Plan* p, const mcl_run_params* runs, int32_t n_runs,
                  const std::vector<mcl::RunDesc>& desc)
{
    struct Item { int32_t cluster; int32_t cls; double work; };
    std::vector<Item> items;
    items.reserve((size_t)p->total);
    int64_t k = 0;
    for (int r = 0; r < n_runs; ++r) {
        int cls = 0;
        while ((128 << cls) < desc[(size_t)r].n_e0) ++cls;
        const double work = (double)(desc[(size_t)r].n_e0 + 1) * (double)runs[r].n_steps;
        for (int s = 0; s < runs[r].sims; ++s, ++k) items.push_back({(int32_t)k, cls, work});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return a.cls != b.cls ? a.cls > b.cls : a.work > b.work;
    });
    std::vector<int32_t> order(items.size());
    p->classes.clear();
    for (size_t i = 0; i < items.size(); ++i) {
        order[i] = items[i].cluster;
        if (i == 0 || items[i].cls != items[i - 1].cls) {
            Plan::Class cl;
            cl.first = (int32_t)i;
            p->classes.push_back(cl);
        }
        ++p->classes.back().count;
    }
    // class capacities from the runs of their clusters
    std::vector<int32_t> c_run((size_t)p->total);
    k = 0;
    for (int r = 0; r < n_runs; ++r)
        for (int s = 0; s < runs[r].sims; ++s, ++k) c_run[(size_t)k] = r;
    for (Plan::Class& cl : p->classes) {
        int max_cells = 1;
        for (int i = cl.first; i < cl.first + cl.count; ++i) {
            const mcl::RunDesc& d = desc[(size_t)c_run[(size_t)order[(size_t)i]]];
            cl.max_e = std::max(cl.max_e, d.n_e0);
            cl.max_h = std::max(cl.max_h, d.n_h);
            max_cells = std::max(max_cells, d.grid * d.grid * d.grid);
            if (keeps_thresholds(c, d)) cl.fixed_thr = 1;
            if (uses_window(c, d)) cl.window = 1;
        }
        cl.h_cap = ((cl.max_h + 3) / 4) * 4 + 4;
        cl.e_cap = std::max(64, ((cl.max_e + 63) / 64) * 64);
        cl.g_cap = ((max_cells + 1 + 3) / 4) * 4;
        const Layout L = layout_for(c, cl.h_cap, cl.e_cap, cl.g_cap);
        cl.lds = L.lds;
        cl.global = L.global;
        cl.e_lds = L.e_lds;
        cl.gstride = L.gstride;
        // (the kernels' dynamic-LDS limit was raised to the plan's largest class)
        int per_cu = 0;
        const int st = choose_waves(c, cl.lds, cl.e_cap, cl.count,
                                    p->classes.size() > 1 ? 4 : 8, &cl.waves, &per_cu, cl.global);
        if (st) return st;
        const double frac = c->persist_frac;
        int grid = (int)std::max(1.0, std::floor((double)std::max(per_cu, 1) * c->n_cu * frac));
        cl.grid = std::min(grid, cl.count);
    }
    // one scratch block for the slices of every global-memory class
    size_t gbytes = 0;
    for (Plan::Class& cl : p->classes) {
        if (!cl.global) continue;
        cl.goff = gbytes;
        gbytes += (size_t)cl.grid * (size_t)cl.gstride;
    }
    const size_t ncls = p->classes.size();
    hipError_t e = mcl::pool_alloc(c, (void**)&p->d_order, sizeof(int32_t) * order.size());
    if (e == hipSuccess)
        e = mcl::pool_alloc(c, (void**)&p->d_tickets, sizeof(uint32_t) * std::max<size_t>(ncls, 1));
    p->class_gbytes = gbytes;      // allocated with the plain launch's slices (make_plan)
    if (e == hipSuccess)
        e = hipMemcpy(p->d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
    for (size_t i = 0; i < ncls && e == hipSuccess; ++i) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) p->join.push_back(ev);
        if (e == hipSuccess && i > 0) {
            hipStream_t s = nullptr;
            e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);   // default priority
            if (e == hipSuccess) p->side.push_back(s);
        }
    }
    if (e != hipSuccess)
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "persistent setup: %s",
                    hipGetErrorString(e));
    return MCL_OK;
}

// Build a plan; run_ids overrides the Philox run index of each run (may be null).
int make_plan(Ctx* c, const mcl_run_params* runs, int32_t n_runs, const int32_t* run_ids,
              uint64_t seed, int32_t n_bins, uint32_t flags, Plan** out)
{
    if (!runs || n_runs <= 0) return fail(c, MCL_EINVAL, "no runs");
    std::vector<mcl::RunDesc> desc((size_t)n_runs);
    int64_t total = 0;
    int max_e = 0, max_h = 0, max_cells = 1, max_steps = 0;
    for (int r = 0; r < n_runs; ++r) {
        mcl_geometry g;
        int st = validate_run(c, runs[r], r, n_bins, g);
        if (st) return st;
        desc[(size_t)r] = make_desc(runs[r], g, run_ids ? run_ids[r] : r, c->holes_per_cell);
        total += runs[r].sims;
        if (runs[r].sims > 0) {
            if (g.n_e0 > max_e) max_e = g.n_e0;
            if (g.n_h > max_h) max_h = g.n_h;
            if (runs[r].n_steps > max_steps) max_steps = runs[r].n_steps;
            const int G = desc[(size_t)r].grid;
            if (G * G * G > max_cells) max_cells = G * G * G;
        }
    }
    if (total <= 0) return fail(c, MCL_EINVAL, "no clusters (sum of sims is 0)");
    if (total > 0x7FFFFFFF) return fail(c, MCL_EINVAL, "too many clusters (%lld)", (long long)total);
    const int h_cap = ((max_h + 3) / 4) * 4 + 4;
    const int e_cap = std::max(64, ((max_e + 63) / 64) * 64);
    const int g_cap = ((max_cells + 1 + 3) / 4) * 4;
    const Layout L = layout_for(c, h_cap, e_cap, g_cap);
    const int lds = L.lds;
    std::vector<int32_t> crun((size_t)total), csim((size_t)total);
    int64_t k = 0;
    for (int r = 0; r < n_runs; ++r)
        for (int s = 0; s < runs[r].sims; ++s, ++k) {
            crun[(size_t)k] = r;
            csim[(size_t)k] = s;
        }
    Plan* p = new (std::nothrow) Plan();
    if (!p) return fail(c, MCL_ENOMEM, "host allocation failed");
    p->ctx = c;
    p->key0 = (uint32_t)seed;
    p->key1 = (uint32_t)(seed >> 32);
    p->total = total;
    p->n_bins = n_bins;
    p->row_len = n_bins > 0 ? n_bins : max_steps;
    p->h_cap = h_cap;
    p->e_cap = e_cap;
    p->g_cap = g_cap;
    p->lds = lds;
    p->max_e = max_e;
    p->max_h = max_h;
    p->flags = flags;
    for (int r = 0; r < n_runs; ++r)
        if (runs[r].sims > 0 && keeps_thresholds(c, desc[(size_t)r])) p->fixed_thr = 1;
    for (int r = 0; r < n_runs; ++r)
        if (runs[r].sims > 0 && uses_window(c, desc[(size_t)r])) p->window = 1;
    p->global = L.global;
    p->e_lds = L.e_lds;
    p->gstride = L.gstride;
    hipError_t e = mcl::pool_alloc(c, (void**)&p->d_runs, sizeof(mcl::RunDesc) * (size_t)n_runs);
    if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&p->d_crun, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&p->d_csim, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess)
        e = mcl::pool_alloc(c, (void**)&p->d_work, sizeof(uint64_t) * 4 * (size_t)total);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_runs, desc.data(), sizeof(mcl::RunDesc) * (size_t)n_runs,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_crun, crun.data(), sizeof(int32_t) * (size_t)total,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_csim, csim.data(), sizeof(int32_t) * (size_t)total,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventCreate(&p->ev0);
    if (e == hipSuccess) e = hipEventCreate(&p->ev1);
    // (LDS limits: the plan's largest need, at most a workgroup's; classes need less)
    if (e == hipSuccess)
        e = mcl::prepare_fixed_dt(mcl::fixed_dt_lds_bytes(h_cap, e_cap, g_cap), c->max_lds);
    if (e == hipSuccess && L.global) e = mcl::prepare_fixed_dt_global(c->max_lds);
    if (e != hipSuccess) {
        plan_free(p, false);
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "plan setup: %s",
                    hipGetErrorString(e));
    }
    int per_cu = 0;
    {
        const int st = choose_waves(c, lds, e_cap, total, 8, &p->waves, &per_cu, L.global);
        if (st) {
            plan_free(p, true);
            return st;
        }
    }
    if (flags & (MCL_F_PERSISTENT | MCL_F_GRAPH)) {
        const int st = build_classes(c, p, runs, n_runs, desc);
        if (st) {
            plan_free(p, true);
            return st;
        }
    }
    // global-memory form: one scratch slice per resident workgroup of the plain
    // launch (its grid), or of the classes, and the plain launch's ticket
    if (L.global) {
        const int64_t resident = (int64_t)std::max(per_cu, 1) * c->n_cu;
        p->ggrid = (int32_t)std::min<int64_t>(total, std::max<int64_t>(resident, 1));
        const size_t plain = (size_t)p->ggrid * (size_t)p->gstride;
        const size_t need = std::max(plain, p->class_gbytes);
        if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&p->d_gscratch, need);
        if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&p->d_gticket, sizeof(uint32_t));
        if (e != hipSuccess) {
            plan_free(p, true);
            return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE,
                        "global-memory scratch (%zu B): %s", need, hipGetErrorString(e));
        }
    }
    *out = p;
    return MCL_OK;
}

mcl::LaunchArgs launch_args(const Plan* p, uint32_t* counts, uint64_t* live)
{
    mcl::LaunchArgs a{};
    a.runs = p->d_runs;
    a.c_run = p->d_crun;
    a.c_sim = p->d_csim;
    a.n_clusters = (int32_t)p->total;
    a.order = nullptr;
    a.n_order = (int32_t)p->total;
    a.queue = nullptr;
    a.key0 = p->key0;
    a.key1 = p->key1;
    a.n_bins = p->n_bins;
    a.row_len = p->row_len;
    a.counts = counts;
    a.live_steps = live;
    a.work = p->d_work;
    a.h_cap = p->h_cap;
    a.e_cap = p->e_cap;
    a.g_cap = p->g_cap;
    a.wave_search_max = (int32_t)p->ctx->wave_search_max;
    a.lookahead_max = (int32_t)p->ctx->lookahead_max;
    a.quiet_window = p->window;
    a.dump.enabled = 0;
    a.fixed_thr = p->fixed_thr;
    return a;
}

// The persistent launch on stream s: reset the tickets, fork one stream per
// class, run the class kernels concurrently, join.
hipError_t enqueue_persistent(Plan* p, uint32_t* counts, uint64_t* live, hipStream_t s)
{
    const int ncls = (int)p->classes.size();
    hipError_t e = mcl::launch_reset_tickets(p->d_tickets, ncls, s);
    if (e == hipSuccess) e = hipEventRecord(p->fork, s);
    for (int k = 0; k < ncls && e == hipSuccess; ++k) {
        const Plan::Class& cl = p->classes[(size_t)k];
        hipStream_t sk = k == 0 ? s : p->side[(size_t)k - 1];
        if (k > 0) e = hipStreamWaitEvent(sk, p->fork, 0);
        mcl::LaunchArgs a = launch_args(p, counts, live);
        a.order = p->d_order + cl.first;
        a.n_order = cl.count;
        a.queue = p->d_tickets + k;
        a.fixed_thr = cl.fixed_thr;
        a.quiet_window = cl.window;
        a.h_cap = cl.h_cap;
        a.e_cap = cl.e_cap;
        a.g_cap = cl.g_cap;
        if (cl.global) {
            a.gscratch = p->d_gscratch + cl.goff;
            a.gstride = cl.gstride;
            a.e_lds = cl.e_lds;
            if (e == hipSuccess) e = mcl::launch_fixed_dt_global(a, cl.grid, cl.lds, cl.waves, sk);
        } else if (e == hipSuccess) {
            e = mcl::launch_fixed_dt(a, cl.grid, cl.lds, cl.waves, sk);
        }
        if (e == hipSuccess && k > 0) e = hipEventRecord(p->join[(size_t)k], sk);
        if (e == hipSuccess && k > 0) e = hipStreamWaitEvent(s, p->join[(size_t)k], 0);
    }
    return e;
}

int plan_launch(Plan* p, uint32_t* counts, uint64_t* live, const mcl::StateDump* dump)
{
    Ctx* c = p->ctx;
    if (!counts) return fail(c, MCL_EINVAL, "counts is null");
    if (p->flags & MCL_F_GRAPH && !dump) {
        if (!p->gexec || p->g_counts != counts || p->g_live != live) {
            if (p->gexec) {
                HIP_TRY(c, hipGraphExecDestroy(p->gexec));
                p->gexec = nullptr;
            }
            hipGraph_t graph = nullptr;
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_persistent(p, counts, live, c->stream);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (graph) hipGraphDestroy(graph);
                return fail(c, MCL_EDEVICE, "graph capture: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2));
            }
            e = hipGraphInstantiate(&p->gexec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(c, MCL_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
            p->g_counts = counts;
            p->g_live = live;
        }
        HIP_TRY(c, hipEventRecord(p->ev0, c->stream));
        HIP_TRY(c, hipGraphLaunch(p->gexec, c->stream));
        HIP_TRY(c, hipEventRecord(p->ev1, c->stream));
        p->timed = true;
        return MCL_OK;
    }
    if (p->flags & MCL_F_PERSISTENT && !dump) {
        HIP_TRY(c, hipEventRecord(p->ev0, c->stream));
        HIP_TRY(c, enqueue_persistent(p, counts, live, c->stream));
        HIP_TRY(c, hipEventRecord(p->ev1, c->stream));
        p->timed = true;
        return MCL_OK;
    }
    mcl::LaunchArgs a = launch_args(p, counts, live);
    if (dump) a.dump = *dump;
    HIP_TRY(c, hipEventRecord(p->ev0, c->stream));
    // (state dumps run one wave per cluster)
    if (p->global) {
        // persistent over the plan's clusters in order, one scratch slice per workgroup
        a.queue = p->d_gticket;
        a.gscratch = p->d_gscratch;
        a.gstride = p->gstride;
        a.e_lds = p->e_lds;
        const int grid = (int)std::min<int64_t>(p->ggrid, p->total);
        HIP_TRY(c, mcl::launch_reset_tickets(p->d_gticket, 1, c->stream));
        HIP_TRY(c, mcl::launch_fixed_dt_global(a, grid, p->lds, dump ? 1 : p->waves, c->stream));
    } else {
        HIP_TRY(c, mcl::launch_fixed_dt(a, (int)p->total, p->lds, dump ? 1 : p->waves, c->stream));
    }
    HIP_TRY(c, hipEventRecord(p->ev1, c->stream));
    p->timed = true;
    return MCL_OK;
}

}  // namespace

extern "C" {

int mcl_abi_version(void) { return MCL_ABI_VERSION; }

int mcl_create(void** ctx_out, int hip_device)
{
    if (!ctx_out) return MCL_EINVAL;
    *ctx_out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MCL_EDEVICE;
    if (hip_device < 0 || hip_device >= n) return MCL_EINVAL;
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return MCL_ENOMEM;
    c->device = hip_device;
    hipDeviceProp_t prop;
    if (hipSetDevice(hip_device) != hipSuccess ||
        hipGetDeviceProperties(&prop, hip_device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MCL_EDEVICE;
    }
    c->stream = c->own_stream;
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = (int)prop.sharedMemPerBlock;
    if (hipEventCreate(&c->na_ev0) != hipSuccess || hipEventCreate(&c->na_ev1) != hipSuccess) {
        if (c->na_ev0) hipEventDestroy(c->na_ev0);
        hipStreamDestroy(c->own_stream);
        delete c;
        return MCL_EDEVICE;
    }
    init_scratch(hip_device);
    *ctx_out = c;
    return MCL_OK;
}

void mcl_destroy(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return;
    if (c->own_stream) {
        hipStreamSynchronize(c->own_stream);
        hipStreamDestroy(c->own_stream);
    }
    if (c->na_ev0) hipEventDestroy(c->na_ev0);
    if (c->na_ev1) hipEventDestroy(c->na_ev1);
    mcl::pool_clear(c);
    delete c;
}

const char* mcl_last_error(const void* ctx)
{
    const Ctx* c = (const Ctx*)ctx;
    return c ? c->err.c_str() : "null context";
}

int mcl_set_stream(void* ctx, void* hip_stream)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    c->stream = hip_stream ? (hipStream_t)hip_stream : c->own_stream;
    return MCL_OK;
}

int mcl_cluster_limits(const void* ctx, int32_t* max_electrons, int32_t* max_holes)
{
    const Ctx* c = (const Ctx*)ctx;
    if (!c || !max_electrons || !max_holes) return MCL_EINVAL;
    // equal electron and hole counts, about 2 holes per grid cell:
    // 16 (h + 4) + 4 (h / 2 + 4) + 16 e <= max_lds
    int n = (c->max_lds - 16 * 4 - 4 * 4) / 34;
    n = (n / 64) * 64;
    *max_electrons = n > 65535 ? 65535 : n;
    *max_holes = n > 65534 ? 65534 : n;
    return MCL_OK;
}

int mcl_tune_set(void* ctx, const char* key, double value)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !key) return MCL_EINVAL;
    if (std::strcmp(key, "holes_per_cell") == 0) {
        if (!(value >= 0.25 && value <= 64.0))
            return fail(c, MCL_EINVAL, "holes_per_cell must be in [0.25, 64] (got %g)", value);
        c->holes_per_cell = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "wave_search_max") == 0) {
        if (!(value >= 0.0 && value <= 64.0))
            return fail(c, MCL_EINVAL, "wave_search_max must be in [0, 64] (got %g)", value);
        c->wave_search_max = std::floor(value);
        return MCL_OK;
    }
    if (std::strcmp(key, "lookahead_max") == 0) {
        if (!(value >= 0.0 && value <= 65535.0))
            return fail(c, MCL_EINVAL, "lookahead_max must be in [0, 65535] (got %g)", value);
        c->lookahead_max = std::floor(value);
        return MCL_OK;
    }
    if (std::strcmp(key, "persist_frac") == 0) {
        if (!(value > 0.0 && value <= 4.0))
            return fail(c, MCL_EINVAL, "persist_frac must be in (0, 4] (got %g)", value);
        c->persist_frac = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "waves_per_cluster") == 0) {
        if (!(value == 0.0 || value == 1.0 || value == 2.0 || value == 4.0 || value == 8.0))
            return fail(c, MCL_EINVAL, "waves_per_cluster must be 0, 1, 2, 4 or 8 (got %g)", value);
        c->waves_per_cluster = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "mlp_fused") == 0) {
        if (!(value == 0.0 || value == 1.0))
            return fail(c, MCL_EINVAL, "mlp_fused must be 0 or 1 (got %g)", value);
        c->mlp_fused = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "global_memory") == 0) {
        if (!(value == 0.0 || value == 1.0))
            return fail(c, MCL_EINVAL, "global_memory must be 0 or 1 (got %g)", value);
        c->global_memory = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "fixed_thresholds") == 0) {
        if (!(value == 0.0 || value == 1.0))
            return fail(c, MCL_EINVAL, "fixed_thresholds must be 0 or 1 (got %g)", value);
        c->fixed_thresholds = value;
        return MCL_OK;
    }
    if (std::strcmp(key, "quiet_window") == 0) {
        if (!(value == 0.0 || value == 1.0))
            return fail(c, MCL_EINVAL, "quiet_window must be 0 or 1 (got %g)", value);
        c->quiet_window = value;
        return MCL_OK;
    }
    return fail(c, MCL_EINVAL, "unknown tuning key '%s'", key);
}

int mcl_tune_get(const void* ctx, const char* key, double* value)
{
    const Ctx* c = (const Ctx*)ctx;
    if (!c || !key || !value) return MCL_EINVAL;
    if (std::strcmp(key, "holes_per_cell") == 0) {
        *value = c->holes_per_cell;
        return MCL_OK;
    }
    if (std::strcmp(key, "persist_frac") == 0) {
        *value = c->persist_frac;
        return MCL_OK;
    }
    if (std::strcmp(key, "lookahead_max") == 0) {
        *value = c->lookahead_max;
        return MCL_OK;
    }
    if (std::strcmp(key, "wave_search_max") == 0) {
        *value = c->wave_search_max;
        return MCL_OK;
    }
    if (std::strcmp(key, "waves_per_cluster") == 0) {
        *value = c->waves_per_cluster;
        return MCL_OK;
    }
    if (std::strcmp(key, "mlp_fused") == 0) {
        *value = c->mlp_fused;
        return MCL_OK;
    }
    if (std::strcmp(key, "global_memory") == 0) {
        *value = c->global_memory;
        return MCL_OK;
    }
    if (std::strcmp(key, "fixed_thresholds") == 0) {
        *value = c->fixed_thresholds;
        return MCL_OK;
    }
    if (std::strcmp(key, "quiet_window") == 0) {
        *value = c->quiet_window;
        return MCL_OK;
    }
    return MCL_EINVAL;
}

int mcl_run_geometry(const mcl_run_params* run, mcl_geometry* out)
{
    if (!run || !out) return MCL_EINVAL;
    geometry(*run, *out);
    return MCL_OK;
}

int mcl_plan_fixed_dt(void* ctx, const mcl_run_params* runs, int32_t n_runs, uint64_t seed,
                      int32_t n_bins, uint32_t flags, void** plan_out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !plan_out) return MCL_EINVAL;
    *plan_out = nullptr;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    Plan* p = nullptr;
    int st = make_plan(c, runs, n_runs, nullptr, seed, n_bins, flags, &p);
    if (st) return st;
    *plan_out = p;
    return MCL_OK;
}

int mcl_plan_info_get(const void* plan, mcl_plan_info* info)
{
    const Plan* p = (const Plan*)plan;
    if (!p || !info) return MCL_EINVAL;
    info->total_clusters = p->total;
    info->n_bins = p->n_bins;
    info->row_len = p->row_len;
    info->lds_bytes = p->lds + mcl::fixed_dt_waves_lds(p->waves);
    info->max_electrons = p->max_e;
    info->max_holes = p->max_h;
    info->grid = p->global ? (int32_t)std::min<int64_t>(p->ggrid, p->total) : (int32_t)p->total;
    info->block = mcl::kWaveHost * p->waves;
    return MCL_OK;
}

int mcl_plan_launch(void* plan, uint32_t* counts_dev, uint64_t* live_steps_dev)
{
    Plan* p = (Plan*)plan;
    if (!p) return MCL_EINVAL;
    return plan_launch(p, counts_dev, live_steps_dev, nullptr);
}

int mcl_plan_last_ms(void* plan, double* ms)
{
    Plan* p = (Plan*)plan;
    if (!p || !ms) return MCL_EINVAL;
    if (!p->timed) return fail(p->ctx, MCL_EINVAL, "plan has not been launched");
    HIP_TRY(p->ctx, hipEventSynchronize(p->ev1));
    float f = 0.0f;
    HIP_TRY(p->ctx, hipEventElapsedTime(&f, p->ev0, p->ev1));
    *ms = (double)f;
    return MCL_OK;
}

int mcl_plan_work(void* plan, mcl_work* out)
{
    Plan* p = (Plan*)plan;
    if (!p || !out) return MCL_EINVAL;
    Ctx* c = p->ctx;
    if (!p->timed) return fail(c, MCL_EINVAL, "plan has not been launched");
    std::vector<uint64_t> w(4 * (size_t)p->total);
    HIP_TRY(c, hipMemcpyAsync(w.data(), p->d_work, sizeof(uint64_t) * w.size(), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::memset(out, 0, sizeof *out);
    for (size_t i = 0; i < (size_t)p->total; ++i) {
        out->live_electron_steps += w[4 * i];
        out->distance_evals += w[4 * i + 1];
        out->recombinations += w[4 * i + 2];
        out->cluster_steps += w[4 * i + 3];
    }
    float f = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&f, p->ev0, p->ev1));
    out->kernel_ms = (double)f;
    return MCL_OK;
}

void mcl_plan_destroy(void* plan) { plan_free((Plan*)plan, false); }

int mcl_sync(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MCL_OK;
}

int mcl_run_lab(void* ctx, const mcl_run_params* runs, int32_t n_runs, const double* obs_times,
                const int32_t* obs_count, uint32_t options, uint64_t seed, int64_t max_steps,
                double* ratio, int32_t* status, mcl_stats* stats)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!runs || !obs_times || !obs_count || !ratio)
        return fail(c, MCL_EINVAL, "runs, obs_times, obs_count and ratio must not be null");
    if (options & ~MCL_LAB_LEGACY) return fail(c, MCL_EINVAL, "unknown option bits 0x%x", options);
    if (n_runs <= 0) return fail(c, MCL_EINVAL, "n_runs must be > 0");
    if (max_steps <= 0) return fail(c, MCL_EINVAL, "max_steps must be > 0");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    std::vector<mcl::RunDesc> desc((size_t)n_runs);
    std::vector<int32_t> offs((size_t)n_runs);
    int64_t total = 0, n_obs_all = 0;
    int max_obs = 0, max_e = 0, max_h = 0;
    for (int r = 0; r < n_runs; ++r) {
        const mcl_run_params& p = runs[r];
        if (!finite_all(p)) return fail(c, MCL_EINVAL, "run %d: non-finite parameter", r);
        if (p.sims < 0) return fail(c, MCL_EINVAL, "run %d: sims must be >= 0", r);
        if (!(p.D >= 0.0)) return fail(c, MCL_EINVAL, "run %d: dose rate D must be >= 0", r);
        if (p.D > 0.0 && !(p.D0 > 0.0)) return fail(c, MCL_EINVAL, "run %d: D0 must be > 0", r);
        if (!(p.k_b > 0.0) || !(p.N_e >= 1.0)) return fail(c, MCL_EINVAL, "run %d: k_b > 0, N_e >= 1", r);
        if (!(p.alpha > 0.0) || !(p.rho_prime > 0.0) || !(p.holes > 0.0) || !(p.boundary_factor > 0.0))
            return fail(c, MCL_EINVAL, "run %d: alpha, rho_prime, holes, boundary_factor must be > 0", r);
        if (p.s < 0.0 || p.b < 0.0) return fail(c, MCL_EINVAL, "run %d: s and b must be >= 0", r);
        const int k = obs_count[r];
        if (k <= 0) return fail(c, MCL_EINVAL, "run %d: needs at least one observation time", r);
        offs[(size_t)r] = (int32_t)n_obs_all;
        for (int j = 0; j < k; ++j) {
            const double o = obs_times[n_obs_all + j];
            if (!std::isfinite(o) || o < 0.0 || (j > 0 && o < obs_times[n_obs_all + j - 1]))
                return fail(c, MCL_EINVAL, "run %d: observation times must be finite, >= 0 and "
                            "ascending", r);
        }
        const double t_end = obs_times[n_obs_all + k - 1];
        if (!((p.T_start + 273.15) > 0.0) || !((p.T_start + 273.15) + p.T_rate * t_end > 0.0))
            return fail(c, MCL_EINVAL, "run %d: temperature must stay > 0 K", r);
        n_obs_all += k;
        max_obs = std::max(max_obs, k);
        if (!(p.holes < 1e6) || !(p.N_e < 1e6) || !(p.boundary_factor < 100.0))
            return fail(c, MCL_ERANGE, "run %d: lab clusters hold at most 8192 electrons and "
                        "16384 holes", r);
        mcl_geometry g;
        if (options & MCL_LAB_LEGACY_GEOMETRY) legacy_geometry(p, g);
        else geometry(p, g);
        if (g.n_e0 < 0 || g.n_h < 0 || g.n_e0 > 8192 || g.n_h > 16384 || p.N_e > 8192)
            return fail(c, MCL_ERANGE, "run %d: lab clusters hold at most 8192 electrons and "
                        "16384 holes", r);
        desc[(size_t)r] = make_desc(p, g, r, c->holes_per_cell);
        total += p.sims;
        if (p.sims > 0) {
            max_e = std::max(max_e, std::max(g.n_e0, (int)p.N_e));
            max_h = std::max(max_h, g.n_h);
        }
    }
    if (total <= 0) return fail(c, MCL_EINVAL, "no clusters (sum of sims is 0)");
    const int e_cap = ((max_e + 64 + 63) / 64) * 64;
    const int h_cap = ((max_h + 2 * max_e + 64 + 3) / 4) * 4;
    const int lds = mcl::lab_lds_bytes(h_cap, e_cap);
    if (lds > c->max_lds)
        return fail(c, MCL_ERANGE, "lab cluster needs %d B of LDS; a workgroup has %d B", lds,
                    c->max_lds);
    std::vector<int32_t> crun((size_t)total), csim((size_t)total);
    {
        int64_t k = 0;
        for (int r = 0; r < n_runs; ++r)
            for (int s2 = 0; s2 < runs[r].sims; ++s2, ++k) {
                crun[(size_t)k] = r;
                csim[(size_t)k] = s2;
            }
    }
    std::vector<void*> tmp;
    auto dmalloc = [&](void** ptr, size_t bytes) {
        const hipError_t e = mcl::pool_alloc(c, ptr, bytes);
        if (e == hipSuccess) tmp.push_back(*ptr);
        return e;
    };
    // idle: the stream was synchronised after the last use (else a device sync first)
    auto release = [&](bool idle) {
        if (!idle) (void)hipDeviceSynchronize();
        for (void* q : tmp) mcl::pool_release(c, q, true);
    };
    mcl::RunDesc* d_runs = nullptr;
    int32_t *d_crun = nullptr, *d_csim = nullptr, *d_off = nullptr, *d_cnt = nullptr,
            *d_status = nullptr;
    double *d_obs = nullptr, *d_ratio = nullptr;
    uint64_t* d_work = nullptr;
    const size_t n_ratio = (size_t)total * (size_t)max_obs;
    hipError_t e = dmalloc((void**)&d_runs, sizeof(mcl::RunDesc) * desc.size());
    if (e == hipSuccess) e = dmalloc((void**)&d_crun, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_csim, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_off, sizeof(int32_t) * (size_t)n_runs);
    if (e == hipSuccess) e = dmalloc((void**)&d_cnt, sizeof(int32_t) * (size_t)n_runs);
    if (e == hipSuccess) e = dmalloc((void**)&d_status, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_obs, sizeof(double) * (size_t)n_obs_all);
    if (e == hipSuccess) e = dmalloc((void**)&d_ratio, sizeof(double) * n_ratio);
    if (e == hipSuccess) e = dmalloc((void**)&d_work, sizeof(uint64_t) * 4 * (size_t)total);
    if (e != hipSuccess) {
        release(true);
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "allocation: %s",
                    hipGetErrorString(e));
    }
    auto h2d = [&](void* d, const void* h, size_t bytes) {
        return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream);
    };
    e = h2d(d_runs, desc.data(), sizeof(mcl::RunDesc) * desc.size());
    if (e == hipSuccess) e = h2d(d_crun, crun.data(), sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = h2d(d_csim, csim.data(), sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = h2d(d_off, offs.data(), sizeof(int32_t) * (size_t)n_runs);
    if (e == hipSuccess) e = h2d(d_cnt, obs_count, sizeof(int32_t) * (size_t)n_runs);
    if (e == hipSuccess) e = h2d(d_obs, obs_times, sizeof(double) * (size_t)n_obs_all);
    if (e == hipSuccess) e = mcl::prepare_lab(lds);
    mcl::LabArgs a{};
    a.runs = d_runs;
    a.c_run = d_crun;
    a.c_sim = d_csim;
    a.n_clusters = (int32_t)total;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.obs = d_obs;
    a.obs_offset = d_off;
    a.obs_count = d_cnt;
    a.max_obs = max_obs;
    a.ratio = d_ratio;
    a.status = d_status;
    a.work = d_work;
    a.max_steps = max_steps;
    a.h_cap = h_cap;
    a.e_cap = e_cap;
    a.options = options;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) e = hipEventRecord(ev0, c->stream);
    if (e == hipSuccess) e = mcl::launch_lab(a, lds, c->stream);
    if (e == hipSuccess) e = hipEventRecord(ev1, c->stream);
    std::vector<int32_t> hs((size_t)total);
    std::vector<uint64_t> hw(4 * (size_t)total);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ratio, d_ratio, sizeof(double) * n_ratio, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hs.data(), d_status, sizeof(int32_t) * (size_t)total,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hw.data(), d_work, sizeof(uint64_t) * 4 * (size_t)total,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ev0, ev1);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    release(e == hipSuccess);
    if (e != hipSuccess)
        return fail(c, MCL_EDEVICE, "lab kernel or copy failed: %s", hipGetErrorString(e));
    if (status) std::memcpy(status, hs.data(), sizeof(int32_t) * (size_t)total);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->kernel_ms = ms;
        stats->clusters = (uint64_t)total;
        for (int64_t k = 0; k < total; ++k) {
            stats->live_electron_steps += hw[4 * (size_t)k];
            stats->recombinations += hw[4 * (size_t)k + 2];
            stats->cluster_steps += hw[4 * (size_t)k + 3];
        }
    }
    for (int64_t k = 0; k < total; ++k) {
        if (hs[(size_t)k] == -1)
            return fail(c, MCL_EOVERFLOW, "cluster %lld (run %d): more than max_steps=%lld steps "
                        "before the last observation", (long long)k, crun[(size_t)k],
                        (long long)max_steps);
        if (hs[(size_t)k] == -2)
            return fail(c, MCL_ERANGE, "cluster %lld (run %d): electrons or holes beyond the lab "
                        "cluster capacity", (long long)k, crun[(size_t)k]);
    }
    return MCL_OK;
}

int mcl_run_event(void* ctx, const mcl_run_params* runs, int32_t n_runs, const double* duration,
                  uint64_t seed, int32_t max_steps, uint32_t flags, uint8_t* lum, double* t_axis,
                  double* e_ratio, int32_t* steps, mcl_stats* stats)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!runs || !duration || !lum || !steps)
        return fail(c, MCL_EINVAL, "runs, duration, lum and steps must not be null");
    if (n_runs <= 0) return fail(c, MCL_EINVAL, "n_runs must be > 0");
    if (max_steps <= 0) return fail(c, MCL_EINVAL, "max_steps must be > 0");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    std::vector<mcl::RunDesc> desc((size_t)n_runs);
    int64_t total = 0;
    int max_e = 0, max_h = 0, max_cells = 1;
    for (int r = 0; r < n_runs; ++r) {
        mcl_geometry g;
        const int st = validate_run(c, runs[r], r, 0, g);
        if (st) return st;
        const double dur = duration[r];
        if (!std::isfinite(dur) || dur < 0.0)
            return fail(c, MCL_EINVAL, "run %d: duration must be finite and >= 0 (got %g)", r, dur);
        if (!((runs[r].T_start + 273.15) + runs[r].T_rate * dur > 0.0))
            return fail(c, MCL_EINVAL, "run %d: temperature must stay > 0 K over the duration", r);
        if (!(runs[r].N_e > 0.0)) return fail(c, MCL_EINVAL, "run %d: N_e must be > 0", r);
        desc[(size_t)r] = make_desc(runs[r], g, r, c->holes_per_cell);
        desc[(size_t)r].duration = dur;
        total += runs[r].sims;
        if (runs[r].sims > 0) {
            max_e = std::max(max_e, g.n_e0);
            max_h = std::max(max_h, g.n_h);
            const int G = desc[(size_t)r].grid;
            max_cells = std::max(max_cells, G * G * G);
        }
    }
    if (total <= 0) return fail(c, MCL_EINVAL, "no clusters (sum of sims is 0)");
    if (total > 0x7FFFFFFF) return fail(c, MCL_EINVAL, "too many clusters");
    const int h_cap = ((max_h + 3) / 4) * 4 + 4;
    const int e_cap = std::max(64, ((max_e + 63) / 64) * 64);
    const int g_cap = ((max_cells + 1 + 3) / 4) * 4;
    // clusters beyond LDS: holes (and electrons, when they do not fit either) in a
    // scratch slice per resident workgroup, persistent over a ticket
    const Layout L = layout_for(c, h_cap, e_cap, g_cap);
    const int lds = L.lds;
    int ggrid = 0;
    if (L.global) {
        int pc = 0;
        const hipError_t oe = mcl::event_global_occupancy(lds, &pc);
        if (oe != hipSuccess)
            return fail(c, MCL_EDEVICE, "occupancy query: %s", hipGetErrorString(oe));
        ggrid = (int)std::min<int64_t>(total, (int64_t)std::max(pc, 1) * c->n_cu);
    }
    std::vector<int32_t> crun((size_t)total), csim((size_t)total);
    {
        int64_t k = 0;
        for (int r = 0; r < n_runs; ++r)
            for (int s2 = 0; s2 < runs[r].sims; ++s2, ++k) {
                crun[(size_t)k] = r;
                csim[(size_t)k] = s2;
            }
    }
    const bool dev = (flags & MCL_F_DEVICE_PTRS) != 0;
    const size_t rowsz = (size_t)total * (size_t)max_steps;
    std::vector<void*> tmp;
    auto dmalloc = [&](void** ptr, size_t bytes) {
        const hipError_t e = mcl::pool_alloc(c, ptr, bytes);
        if (e == hipSuccess) tmp.push_back(*ptr);
        return e;
    };
    // idle: the stream was synchronised after the last use (else a device sync first)
    auto release = [&](bool idle) {
        if (!idle) (void)hipDeviceSynchronize();
        for (void* q : tmp) mcl::pool_release(c, q, true);
    };
    mcl::RunDesc* d_runs = nullptr;
    int32_t *d_crun = nullptr, *d_csim = nullptr, *d_steps = nullptr;
    uint64_t* d_work = nullptr;
    uint8_t* d_lum = dev ? lum : nullptr;
    double* d_t = dev ? t_axis : nullptr;
    double* d_er = dev ? e_ratio : nullptr;
    hipError_t e = dmalloc((void**)&d_runs, sizeof(mcl::RunDesc) * desc.size());
    if (e == hipSuccess) e = dmalloc((void**)&d_crun, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_csim, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_steps, sizeof(int32_t) * (size_t)total);
    if (e == hipSuccess) e = dmalloc((void**)&d_work, sizeof(uint64_t) * 4 * (size_t)total);
    if (e == hipSuccess && !dev) e = dmalloc((void**)&d_lum, rowsz);
    if (e == hipSuccess && !dev && t_axis) e = dmalloc((void**)&d_t, sizeof(double) * rowsz);
    if (e == hipSuccess && !dev && e_ratio) e = dmalloc((void**)&d_er, sizeof(double) * rowsz);
    unsigned char* d_gscratch = nullptr;
    uint32_t* d_ticket = nullptr;
    if (e == hipSuccess && L.global)
        e = dmalloc((void**)&d_gscratch, (size_t)ggrid * (size_t)L.gstride);
    if (e == hipSuccess && L.global) e = dmalloc((void**)&d_ticket, sizeof(uint32_t));
    if (e != hipSuccess) {
        release(true);
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "allocation: %s",
                    hipGetErrorString(e));
    }
    e = hipMemcpyAsync(d_runs, desc.data(), sizeof(mcl::RunDesc) * desc.size(),
                       hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_crun, crun.data(), sizeof(int32_t) * (size_t)total,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_csim, csim.data(), sizeof(int32_t) * (size_t)total,
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = mcl::prepare_event(std::min(c->max_lds, std::max(lds, 1)));
    mcl::EventArgs a{};
    a.runs = d_runs;
    a.c_run = d_crun;
    a.c_sim = d_csim;
    a.n_clusters = (int32_t)total;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.max_steps = max_steps;
    a.lum = d_lum;
    a.t_axis = d_t;
    a.e_ratio = d_er;
    a.steps = d_steps;
    a.work = d_work;
    a.h_cap = h_cap;
    a.e_cap = e_cap;
    a.g_cap = g_cap;
    a.wave_search_max = (int32_t)c->wave_search_max;
    a.queue = d_ticket;
    a.gscratch = d_gscratch;
    a.gstride = L.gstride;
    a.e_lds = L.e_lds;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) e = hipEventRecord(ev0, c->stream);
    if (e == hipSuccess && L.global) e = mcl::launch_reset_tickets(d_ticket, 1, c->stream);
    if (e == hipSuccess)
        e = L.global ? mcl::launch_event_global(a, ggrid, lds, c->stream)
                     : mcl::launch_event(a, lds, c->stream);
    if (e == hipSuccess) e = hipEventRecord(ev1, c->stream);
    std::vector<int32_t> hs((size_t)total);
    std::vector<uint64_t> hw(4 * (size_t)total);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hs.data(), d_steps, sizeof(int32_t) * (size_t)total,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hw.data(), d_work, sizeof(uint64_t) * 4 * (size_t)total,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !dev) e = hipMemcpyAsync(lum, d_lum, rowsz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !dev && t_axis)
        e = hipMemcpyAsync(t_axis, d_t, sizeof(double) * rowsz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !dev && e_ratio)
        e = hipMemcpyAsync(e_ratio, d_er, sizeof(double) * rowsz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ev0, ev1);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    release(e == hipSuccess);
    if (e != hipSuccess) return fail(c, MCL_EDEVICE, "event kernel or copy failed: %s",
                                     hipGetErrorString(e));
    std::memcpy(steps, hs.data(), sizeof(int32_t) * (size_t)total);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->kernel_ms = ms;
        stats->clusters = (uint64_t)total;
        for (int64_t k = 0; k < total; ++k) {
            stats->live_electron_steps += hw[4 * (size_t)k];
            stats->recombinations += hw[4 * (size_t)k + 2];
            stats->cluster_steps += hw[4 * (size_t)k + 3];
        }
    }
    for (int64_t k = 0; k < total; ++k) {
        if (hs[(size_t)k] == mcl::kEventFillUnsupported)
            return fail(c, MCL_EINVAL, "cluster %lld: a filling event before the end of the run "
                        "(dose filling is not part of the event kernel)", (long long)k);
    }
    for (int64_t k = 0; k < total; ++k) {
        if (hs[(size_t)k] < 0)
            return fail(c, MCL_EOVERFLOW, "cluster %lld (run %d, sim %d) needs more than "
                        "max_steps=%d steps", (long long)k, crun[(size_t)k], csim[(size_t)k],
                        max_steps);
    }
    return MCL_OK;
}

int mcl_run_fixed_dt(void* ctx, const mcl_run_params* runs, int32_t n_runs, uint64_t seed,
                     int32_t n_bins, uint32_t flags, uint32_t* counts, uint64_t* live_steps,
                     mcl_stats* stats)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!counts) return fail(c, MCL_EINVAL, "counts is null");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    Plan* p = nullptr;
    int st = make_plan(c, runs, n_runs, nullptr, seed, n_bins, flags, &p);
    if (st) return st;
    const bool dev = (flags & MCL_F_DEVICE_PTRS) != 0;
    const size_t n_counts = (size_t)p->total * (size_t)p->row_len;
    uint32_t* d_counts = dev ? counts : nullptr;
    uint64_t* d_live = nullptr;
    hipError_t e = hipSuccess;
    if (!dev) e = mcl::pool_alloc(c, (void**)&d_counts, sizeof(uint32_t) * n_counts);
    if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&d_live, sizeof(uint64_t) * (size_t)p->total);
    if (e != hipSuccess) {
        if (!dev) mcl::pool_release(c, d_counts, true);
        plan_free(p, true);
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "output allocation: %s",
                    hipGetErrorString(e));
    }
    st = plan_launch(p, d_counts, d_live, nullptr);
    std::vector<uint64_t> live((size_t)p->total);
    if (st == MCL_OK) {
        e = hipMemcpyAsync(live.data(), d_live, sizeof(uint64_t) * (size_t)p->total,
                           hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && !dev)
            e = hipMemcpyAsync(counts, d_counts, sizeof(uint32_t) * n_counts, hipMemcpyDeviceToHost,
                               c->stream);
        if (e == hipSuccess && live_steps)
            e = dev ? hipMemcpyAsync(live_steps, d_live, sizeof(uint64_t) * (size_t)p->total,
                                     hipMemcpyDeviceToDevice, c->stream)
                    : hipSuccess;
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            st = fail(c, MCL_EDEVICE, "kernel or copy failed: %s", hipGetErrorString(e));
    }
    if (st == MCL_OK && live_steps && !dev)
        std::memcpy(live_steps, live.data(), sizeof(uint64_t) * (size_t)p->total);
    if (st == MCL_OK && stats) {
        std::memset(stats, 0, sizeof *stats);
        double ms = 0.0;
        mcl_plan_last_ms(p, &ms);
        stats->kernel_ms = ms;
        stats->clusters = (uint64_t)p->total;
        mcl_work w;
        if (mcl_plan_work(p, &w) == MCL_OK) {
            stats->live_electron_steps = w.live_electron_steps;
            stats->recombinations = w.recombinations;
            stats->cluster_steps = w.cluster_steps;
        }
    }
    // after a failed launch or copy kernels may still run: release after a device sync
    if (e != hipSuccess || st != MCL_OK) (void)hipDeviceSynchronize();
    if (!dev) mcl::pool_release(c, d_counts, true);
    mcl::pool_release(c, d_live, true);
    plan_free(p, true);
    return st;
}

int mcl_run_fixed_dt_noalloc(void* ctx, const mcl_run_params* runs, int32_t n_runs,
                             uint64_t seed, int32_t n_bins, uint32_t* counts_dev,
                             uint64_t* live_steps_dev, mcl_stats* stats)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!counts_dev) return fail(c, MCL_EINVAL, "counts is null");
    if (!runs || n_runs <= 0) return fail(c, MCL_EINVAL, "no runs");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    hipError_t e = hipSuccess;
    unsigned char* const base = scratch_for(c->device, &e);
    if (!base)
        return fail(c, MCL_EDEVICE, "no scratch address for device %d: %s", c->device,
                    hipGetErrorString(e));
    // sizes, as make_plan derives them
    int64_t total = 0;
    int max_e = 0, max_h = 0, max_cells = 1, max_steps = 0;
    for (int r = 0; r < n_runs; ++r) {
        mcl_geometry g;
        const int st = validate_run(c, runs[r], r, n_bins, g);
        if (st) return st;
        total += runs[r].sims;
        if (runs[r].sims > 0) {
            max_e = std::max(max_e, g.n_e0);
            max_h = std::max(max_h, g.n_h);
            max_steps = std::max(max_steps, (int)runs[r].n_steps);
            const int G = mcl::grid_dim(g.n_h, c->holes_per_cell);
            max_cells = std::max(max_cells, G * G * G);
        }
    }
    if (total <= 0) return fail(c, MCL_EINVAL, "no clusters (sum of sims is 0)");
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t n_cl = (size_t)total;
    const size_t off_crun = up(sizeof(mcl::RunDesc) * (size_t)n_runs);
    const size_t off_csim = off_crun + up(sizeof(int32_t) * n_cl);
    const size_t off_work = off_csim + up(sizeof(int32_t) * n_cl);
    const size_t need = off_work + sizeof(uint64_t) * 4 * n_cl;
    if (need > (size_t)MCL_NOALLOC_SCRATCH_BYTES)
        return fail(c, MCL_ERANGE,
                    "%d runs / %lld clusters need %zu B of scratch; the allocation-free entry "
                    "point has %u B (use mcl_run_fixed_dt)",
                    n_runs, (long long)total, need, (unsigned)MCL_NOALLOC_SCRATCH_BYTES);
    const int h_cap = ((max_h + 3) / 4) * 4 + 4;
    const int e_cap = std::max(64, ((max_e + 63) / 64) * 64);
    const int g_cap = ((max_cells + 1 + 3) / 4) * 4;
    const int lds = mcl::fixed_dt_lds_bytes(h_cap, e_cap, g_cap);
    if (lds > c->max_lds)
        return fail(c, MCL_ERANGE,
                    "cluster of %d electrons / %d holes needs %d B of LDS; a workgroup has %d B "
                    "(larger clusters run in global memory: use mcl_run_fixed_dt)",
                    max_e, max_h, lds, c->max_lds);
    // descriptors and cluster map through a fixed host staging block, one chunk
    // at a time (each copied and waited for before the block is refilled)
    alignas(64) unsigned char stage[16384];
    int32_t fixed_thr = 0, window = 0;
    auto upload = [&](size_t dst_off, size_t bytes) -> hipError_t {
        hipError_t r = hipMemcpyAsync(base + dst_off, stage, bytes, hipMemcpyHostToDevice, c->stream);
        if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
        return r;
    };
    {
        const int per = (int)(sizeof stage / sizeof(mcl::RunDesc));
        mcl::RunDesc* d = (mcl::RunDesc*)stage;
        for (int r0 = 0; r0 < n_runs && e == hipSuccess; r0 += per) {
            const int k = std::min(per, n_runs - r0);
            for (int j = 0; j < k; ++j) {
                mcl_geometry g;
                geometry(runs[r0 + j], g);
                d[j] = make_desc(runs[r0 + j], g, r0 + j, c->holes_per_cell);
                if (runs[r0 + j].sims > 0 && keeps_thresholds(c, d[j])) fixed_thr = 1;
                if (runs[r0 + j].sims > 0 && uses_window(c, d[j])) window = 1;
            }
            e = upload(sizeof(mcl::RunDesc) * (size_t)r0, sizeof(mcl::RunDesc) * (size_t)k);
        }
    }
    {
        const size_t per = sizeof stage / (2 * sizeof(int32_t));
        int32_t* cr = (int32_t*)stage;
        int r = 0, s = 0;          // run and sim of the next cluster
        for (size_t k0 = 0; k0 < n_cl && e == hipSuccess; k0 += per) {
            const size_t k = std::min(per, n_cl - k0);
            int32_t* cs = cr + k;
            for (size_t j = 0; j < k; ++j) {
                while (s >= runs[r].sims) {
                    ++r;
                    s = 0;
                }
                cr[j] = r;
                cs[j] = s++;
            }
            e = upload(off_crun + sizeof(int32_t) * k0, sizeof(int32_t) * k);
            if (e == hipSuccess)
                e = hipMemcpyAsync(base + off_csim + sizeof(int32_t) * k0, cs, sizeof(int32_t) * k,
                                   hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
    }
    if (e == hipSuccess) e = mcl::prepare_fixed_dt(lds, c->max_lds);
    if (e != hipSuccess)
        return fail(c, MCL_EDEVICE, "noalloc setup: %s", hipGetErrorString(e));
    int waves = 1, per_cu = 0;
    int st = choose_waves(c, lds, e_cap, total, 8, &waves, &per_cu, false);
    if (st) return st;
    mcl::LaunchArgs a{};
    a.runs = (const mcl::RunDesc*)base;
    a.c_run = (const int32_t*)(base + off_crun);
    a.c_sim = (const int32_t*)(base + off_csim);
    a.n_clusters = (int32_t)total;
    a.order = nullptr;
    a.n_order = (int32_t)total;
    a.queue = nullptr;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.n_bins = n_bins;
    a.row_len = n_bins > 0 ? n_bins : max_steps;
    a.counts = counts_dev;
    a.live_steps = live_steps_dev;
    a.work = (uint64_t*)(base + off_work);
    a.h_cap = h_cap;
    a.e_cap = e_cap;
    a.g_cap = g_cap;
    a.wave_search_max = (int32_t)c->wave_search_max;
    a.lookahead_max = (int32_t)c->lookahead_max;
    a.quiet_window = window;
    a.dump.enabled = 0;
    a.fixed_thr = fixed_thr;
    e = hipEventRecord(c->na_ev0, c->stream);
    if (e == hipSuccess) e = mcl::launch_fixed_dt(a, (int)total, lds, waves, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->na_ev1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, MCL_EDEVICE, "noalloc launch: %s", hipGetErrorString(e));
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->na_ev0, c->na_ev1) == hipSuccess) stats->kernel_ms = ms;
        stats->clusters = (uint64_t)total;
        // the work counters back through the same staging block
        const size_t per = sizeof stage / (4 * sizeof(uint64_t));
        const uint64_t* w = (const uint64_t*)stage;
        for (size_t k0 = 0; k0 < n_cl && e == hipSuccess; k0 += per) {
            const size_t k = std::min(per, n_cl - k0);
            e = hipMemcpyAsync(stage, base + off_work + 4 * sizeof(uint64_t) * k0,
                               4 * sizeof(uint64_t) * k, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            for (size_t j = 0; j < k && e == hipSuccess; ++j) {
                stats->live_electron_steps += w[4 * j];
                stats->recombinations += w[4 * j + 2];
                stats->cluster_steps += w[4 * j + 3];
            }
        }
        if (e != hipSuccess) return fail(c, MCL_EDEVICE, "noalloc stats: %s", hipGetErrorString(e));
    }
    return MCL_OK;
}

int mcl_run_cluster(void* ctx, const mcl_run_params* run, int32_t run_index, int32_t sim,
                    uint64_t seed, int32_t n_bins, uint32_t* counts, uint64_t* live_steps)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!run || !counts) return fail(c, MCL_EINVAL, "null run or counts");
    if (run_index < 0 || sim < 0) return fail(c, MCL_EINVAL, "run_index and sim must be >= 0");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    mcl_run_params one = *run;
    one.sims = sim + 1;   // validated and sized as clusters 0..sim; only `sim` runs
    Plan* p = nullptr;
    int st = make_plan(c, &one, 1, &run_index, seed, n_bins, 0, &p);
    if (st) return st;
    const size_t row = (size_t)p->row_len;
    uint32_t* d_counts = nullptr;
    uint64_t* d_live = nullptr;
    const int32_t zero = 0;
    hipError_t e = hipMemcpyAsync(p->d_crun, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_csim, &sim, sizeof sim, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);   // host values leave scope
    p->total = 1;
    if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&d_counts, sizeof(uint32_t) * row);
    if (e == hipSuccess) e = mcl::pool_alloc(c, (void**)&d_live, sizeof(uint64_t));
    if (e != hipSuccess) {
        mcl::pool_release(c, d_counts, true);
        plan_free(p, true);
        return fail(c, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "run_cluster setup: %s",
                    hipGetErrorString(e));
    }
    st = plan_launch(p, d_counts, d_live, nullptr);
    uint64_t live = 0;
    if (st == MCL_OK) {
        e = hipMemcpyAsync(counts, d_counts, sizeof(uint32_t) * row, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&live, d_live, sizeof live, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = fail(c, MCL_EDEVICE, "run_cluster: %s", hipGetErrorString(e));
    }
    if (st == MCL_OK && live_steps) *live_steps = live;
    if (e != hipSuccess || st != MCL_OK) (void)hipDeviceSynchronize();
    mcl::pool_release(c, d_counts, true);
    mcl::pool_release(c, d_live, true);
    plan_free(p, true);
    return st;
}

int mcl_cluster_state(void* ctx, const mcl_run_params* run, int32_t run_index, int32_t sim,
                      uint64_t seed, int32_t max_steps, float* e_xyz, float* h_xyz,
                      uint8_t* e_live, uint8_t* h_alive, uint32_t* nn, float* w)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return MCL_EINVAL;
    if (!run || !e_xyz || !h_xyz || !e_live || !h_alive || !nn || !w)
        return fail(c, MCL_EINVAL, "null output pointer");
    if (run_index < 0 || sim < 0 || max_steps < 0)
        return fail(c, MCL_EINVAL, "run_index, sim and max_steps must be >= 0");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MCL_EDEVICE, "hipSetDevice failed");
    mcl_run_params one = *run;
    one.sims = sim + 1;   // clusters 0..sim; only the last one is dumped
    Plan* p = nullptr;
    int st = make_plan(c, &one, 1, &run_index, seed, 1, 0, &p);
    if (st) return st;
    mcl_geometry g;
    geometry(one, g);
    const size_t ne = (size_t)(g.n_e0 > 0 ? g.n_e0 : 0), nh = (size_t)(g.n_h > 0 ? g.n_h : 0);
    // only the requested sim is simulated: a one-cluster map pointing at it
    const int32_t zero = 0;
    hipError_t e = hipMemcpyAsync(p->d_crun, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->d_csim, &sim, sizeof sim, hipMemcpyHostToDevice, c->stream);
    p->total = 1;
    char* buf = nullptr;
    const size_t bytes = ne * 12 + nh * 12 + ne + nh + ne * 4 + ne * 4 + 64 + 64;
    if (e == hipSuccess) e = hipMalloc((void**)&buf, bytes);
    if (e != hipSuccess) {
        plan_free(p, false);
        return fail(c, MCL_EDEVICE, "cluster_state setup: %s", hipGetErrorString(e));
    }
    mcl::StateDump d{};
    size_t off = 0;
    d.e_xyz = (float*)(buf + off); off += ne * 12;
    d.h_xyz = (float*)(buf + off); off += nh * 12;
    d.nn = (uint32_t*)(buf + off); off += ne * 4;
    d.w = (float*)(buf + off); off += ne * 4;
    d.e_live = (uint8_t*)(buf + off); off += ne;
    d.h_alive = (uint8_t*)(buf + off); off += nh;
    uint32_t* d_counts = (uint32_t*)(buf + off + (64 - (off % 64)) % 64);
    d.max_steps = max_steps;
    d.enabled = 1;
    st = plan_launch(p, d_counts, nullptr, &d);
    if (st == MCL_OK) {
        e = hipMemcpyAsync(e_xyz, d.e_xyz, ne * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_xyz, d.h_xyz, nh * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(nn, d.nn, ne * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w, d.w, ne * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(e_live, d.e_live, ne, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_alive, d.h_alive, nh, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = fail(c, MCL_EDEVICE, "cluster_state: %s", hipGetErrorString(e));
    }
    hipFree(buf);
    plan_free(p, false);
    return st;
}

}  // extern "C"
