// mc_common.h — device pieces shared by the fixed-dt and event-mode kernels:
// coordinates from Philox words, the hole cell grid, nearest-hole searches (lane
// private, wave cooperative, full scan; every one returns the first-index argmin
// of the float32 squared distance over live holes, engine.py:113-119,168-175)
// and the per-cluster setup (Box.seed / Box._rebuild, engine.py:113-129).
#pragma once

#include "mcl_device.h"
#include "mcl_internal.h"
#include "../../include/mcl.h"

namespace mcl {

__device__ __forceinline__ float unit_coord(uint32_t word, double side)
{
    return (float)(((double)word * 0x1p-32) * side);
}

__device__ __forceinline__ int cell_of(float v, float inv_cell, int G)
{
    const int c = (int)(v * inv_cell);   // v >= 0
    return c < G ? c : G - 1;
}

__device__ __forceinline__ uint64_t hole_key(float d2, uint32_t meta, uint32_t slot)
{
    // (d2 bits, original index, slot): the min is the first-index argmin of d2
    return ((uint64_t)__float_as_uint(d2) << 32) | ((meta & 0xFFFFu) << 16) | slot;
}

__device__ __forceinline__ float sq_dist(const float4& h, float x, float y, float z)
{
    const float dx = h.x - x, dy = h.y - y, dz = h.z - z;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ uint64_t key_of(const float4& h, float x, float y, float z, int slot)
{
    return hole_key(sq_dist(h, x, y, z), __float_as_uint(h.w), (uint32_t)slot);
}

// Distance from v to the nearest face of cells [lo, hi] that has cells beyond it.
__device__ __forceinline__ float face_gap(float v, int lo, int hi, int G, float cell)
{
    float g = __builtin_inff();
    if (lo > 0) g = v - (float)lo * cell;
    if (hi < G - 1) g = fminf(g, (float)(hi + 1) * cell - v);
    return g;
}

// Is the block minimum d2 provably below every hole outside the block?
__device__ __forceinline__ bool block_conclusive(float best_d2, float gap, float cell)
{
    const float m = gap - 1e-4f * cell;
    return m > 0.0f && best_d2 < m * m * (1.0f - 1e-5f);
}

struct Grid {
    const float4* hole;
    const uint32_t* cstart;
    int G;
    float cell, inv_cell;
    int n_h;
};

struct ElectronGen {
    uint32_t run, sim_tag, k0, k1;
    double side_nm;
    __device__ __forceinline__ void xyz(uint32_t e, float& x, float& y, float& z) const
    {
        const u32x4 o = philox10(0u, e, run, sim_tag, k0, k1);
        x = unit_coord(o.x, side_nm);
        y = unit_coord(o.y, side_nm);
        z = unit_coord(o.z, side_nm);
    }
};

constexpr int kChunk = 8;   // hole records fetched per LDS round trip in a search

// Lane-private search of the (2 RAD + 1)^3 cell block around (x, y, z).
template <int RAD>
__device__ __forceinline__ uint64_t block_search_lane(const Grid& g, float x, float y, float z,
                                                      bool& conclusive, uint64_t& evals)
{
    constexpr int W = 2 * RAD + 1;
    const int G = g.G;
    const int cx = cell_of(x, g.inv_cell, G), cy = cell_of(y, g.inv_cell, G),
              cz = cell_of(z, g.inv_cell, G);
    const int x0 = max(cx - RAD, 0), x1 = min(cx + RAD, G - 1);
    uint64_t key = ~0ull;
#pragma unroll
    for (int k = 0; k < W * W; ++k) {
        const int zz = cz + k / W - RAD, yy = cy + k % W - RAD;
        int s = 0, e = 0;
        if ((unsigned)zz < (unsigned)G && (unsigned)yy < (unsigned)G) {
            const int base = (zz * G + yy) * G;
            s = (int)g.cstart[base + x0];
            e = (int)g.cstart[base + x1 + 1];
        }
        evals += (uint64_t)(e - s);
        for (int i0 = s; i0 < e; i0 += kChunk) {
            float4 hv[kChunk];
#pragma unroll
            for (int t = 0; t < kChunk; ++t) hv[t] = g.hole[min(i0 + t, e - 1)];
#pragma unroll
            for (int t = 0; t < kChunk; ++t) {
                const uint64_t kk = key_of(hv[t], x, y, z, min(i0 + t, e - 1));
                key = kk < key ? kk : key;
            }
        }
    }
    const int y0 = max(cy - RAD, 0), y1 = min(cy + RAD, G - 1);
    const int z0 = max(cz - RAD, 0), z1 = min(cz + RAD, G - 1);
    const bool whole = x0 == 0 && y0 == 0 && z0 == 0 && x1 == G - 1 && y1 == G - 1 && z1 == G - 1;
    const float gap = fminf(face_gap(x, x0, x1, G, g.cell),
                            fminf(face_gap(y, y0, y1, G, g.cell), face_gap(z, z0, z1, G, g.cell)));
    conclusive = whole || block_conclusive(__uint_as_float((uint32_t)(key >> 32)), gap, g.cell);
    return key;
}

// 3x3x3 cell block around (x, y, z): cell bounds, clamped to the grid.
struct Block {
    int x0, x1, y0, y1, z0, z1;
    __device__ __forceinline__ Block(const Grid& g, float x, float y, float z)
    {
        const int G = g.G;
        const int cx = cell_of(x, g.inv_cell, G), cy = cell_of(y, g.inv_cell, G),
                  cz = cell_of(z, g.inv_cell, G);
        x0 = max(cx - 1, 0);
        x1 = min(cx + 1, G - 1);
        y0 = max(cy - 1, 0);
        y1 = min(cy + 1, G - 1);
        z0 = max(cz - 1, 0);
        z1 = min(cz + 1, G - 1);
    }
    // Is the block minimum d2 provably the nearest live hole of (x, y, z)?
    __device__ __forceinline__ bool conclusive(const Grid& g, float x, float y, float z,
                                               float best_d2) const
    {
        const int G = g.G;
        const bool whole = x0 == 0 && y0 == 0 && z0 == 0 && x1 == G - 1 && y1 == G - 1 &&
                           z1 == G - 1;
        const float gap = fminf(face_gap(x, x0, x1, G, g.cell),
                                fminf(face_gap(y, y0, y1, G, g.cell), face_gap(z, z0, z1, G, g.cell)));
        return whole || block_conclusive(best_d2, gap, g.cell);
    }
};

// Wave-cooperative search for one electron (x, y, z wave-uniform).  The 3x3x3
// block is at most 9 contiguous slot ranges (one per (z, y) row of cells); lane
// l takes range l / 7 at offsets l % 7, l % 7 + 7, ... so every lane finds its
// slots from two cell starts, with no cross-lane bookkeeping, then a wave min.
// Ranges longer than 7 holes cost one more pass each 7.
__device__ __forceinline__ uint64_t block_search_wave(const Grid& g, float x, float y, float z,
                                                      int lane, bool& conclusive, uint64_t& evals)
{
    constexpr int kPer = 7;                          // lanes per range (9 x 7 = 63 lanes)
    const int G = g.G;
    const int cx = cell_of(x, g.inv_cell, G), cy = cell_of(y, g.inv_cell, G),
              cz = cell_of(z, g.inv_cell, G);
    const int x0 = max(cx - 1, 0), x1 = min(cx + 1, G - 1);
    const int y0 = max(cy - 1, 0), y1 = min(cy + 1, G - 1);
    const int z0 = max(cz - 1, 0), z1 = min(cz + 1, G - 1);
    const int j = lane / kPer, o = lane - j * kPer;
    // (z, y) row j of the unclamped block; rows outside the grid are empty
    const int jz = j / 3, zz = cz - 1 + jz, yy = cy - 1 + (j - 3 * jz);
    int s = 0, e = 0;
    if (j < 9 && (unsigned)zz < (unsigned)G && (unsigned)yy < (unsigned)G) {
        const int base = (zz * G + yy) * G;
        s = (int)g.cstart[base + x0];
        e = (int)g.cstart[base + x1 + 1];
    }
    if (o == 0) evals += (uint64_t)(e - s);          // summed over lanes at the end
    uint64_t key = ~0ull;
    for (int slot = s + o;; slot += kPer) {
        const bool v = slot < e;
        if (ballot(v) == 0) break;
        if (v) {
            const uint64_t k = key_of(g.hole[slot], x, y, z, slot);
            key = k < key ? k : key;
        }
    }
    key = wave_min_u64(key);
    const bool whole = x0 == 0 && y0 == 0 && z0 == 0 && x1 == G - 1 && y1 == G - 1 && z1 == G - 1;
    const float gap = fminf(face_gap(x, x0, x1, G, g.cell),
                            fminf(face_gap(y, y0, y1, G, g.cell), face_gap(z, z0, z1, G, g.cell)));
    conclusive = whole || block_conclusive(__uint_as_float((uint32_t)(key >> 32)), gap, g.cell);
    return key;
}

// Wave-cooperative 3x3x3 block searches for up to kSearchBatch (four) electrons at
// once, one 16-lane row each (x, y, z: the row's electron, the same in all its
// lanes).  A block is at most 9 contiguous slot ranges (one per (z, y) row of
// cells); lane q of the row takes offsets q, q + 16, ... of every range, so its 9
// range loads are independent; ranges longer than 16 holes take further passes.
// The row minimum is four DPP steps.  Returns the row's block minimum key in every
// lane of the row (~0 for rows at or beyond nb).
constexpr int kSearchBatch = 4;
__device__ __forceinline__ uint64_t block_search_rows(const Grid& g, int nb, float x, float y,
                                                      float z, int lane, uint64_t& evals)
{
    const int G = g.G;
    const int q = lane & 15;
    const bool act = (lane >> 4) < nb;
    int s[9], e[9];
    int longest = 0;
    if (act) {
        const Block k(g, x, y, z);
        const int cz = cell_of(z, g.inv_cell, G), cy = cell_of(y, g.inv_cell, G);
        int total = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            // (z, y) row r of the unclamped block; rows outside the grid are empty
            const int zz = cz - 1 + r / 3, yy = cy - 1 + r % 3;
            s[r] = 0;
            e[r] = 0;
            if ((unsigned)zz < (unsigned)G && (unsigned)yy < (unsigned)G) {
                const int base = (zz * G + yy) * G;
                s[r] = (int)g.cstart[base + k.x0];
                e[r] = (int)g.cstart[base + k.x1 + 1];
            }
            total += e[r] - s[r];
            longest = max(longest, e[r] - s[r]);
        }
        if (q == 0) evals += (uint64_t)total;      // summed over lanes at the end
    } else {
#pragma unroll
        for (int r = 0; r < 9; ++r) s[r] = e[r] = 0;
    }
    uint64_t key = ~0ull;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int slot = s[r] + q;
        if (slot < e[r]) {
            const uint64_t k = key_of(g.hole[slot], x, y, z, slot);
            key = k < key ? k : key;
        }
    }
    if (ballot(longest > 16) != 0) {
        for (int r = 0; r < 9; ++r) {
            for (int slot = s[r] + q + 16; slot < e[r]; slot += 16) {
                const uint64_t k = key_of(g.hole[slot], x, y, z, slot);
                key = k < key ? k : key;
            }
        }
    }
    return row_min_u64(key);
}

// The whole wave scans every hole for one electron (x, y, z wave-uniform).
__device__ __forceinline__ uint64_t full_scan_wave(const Grid& g, float x, float y, float z,
                                                   int lane, uint64_t& evals)
{
    if (lane == 0) evals += (uint64_t)g.n_h;   // evaluations are summed over lanes
    uint64_t key = ~0ull;
    for (int i = lane; i < g.n_h; i += kWave) {
        const uint64_t k = key_of(g.hole[i], x, y, z, i);
        key = k < key ? k : key;
    }
    return wave_min_u64(key);
}

// Nearest live hole for every active lane's electron: 3x3x3 block search per
// lane, then one wave-wide full scan per lane whose block was inconclusive.  (A
// lane-private 5x5x5 level in between measured slower: it is divergent, so one
// open lane makes the wave pay 25 serialized LDS round trips.)
__device__ __forceinline__ uint64_t nn_resolve(const Grid& g, bool act, float x, float y, float z,
                                               int lane, uint64_t& evals)
{
    bool ok = true;
    uint64_t key = ~0ull;
    if (act) key = block_search_lane<1>(g, x, y, z, ok, evals);
    uint64_t fb = ballot(act && !ok);
    while (fb != 0) {
        const int l = __builtin_ctzll(fb);
        fb &= fb - 1;
        const float fx = lane_value(x, l), fy = lane_value(y, l), fz = lane_value(z, l);
        const uint64_t k = full_scan_wave(g, fx, fy, fz, lane, evals);
        if (lane == l) key = k;
    }
    return key;
}

// Inclusive prefix sum of a float64 value over the 64 lanes (all active), in a
// fixed order the CPU replay repeats: within each 16-lane row, steps of 1, 2, 4
// and 8 lanes (v[l] += v[l - o] for l % 16 >= o, by DPP row shifts); then rows
// 1, 2, 3 add t0, t0 + t1 and (t0 + t1) + t2, the totals t_r of the rows before.
template <int kCtrl>
__device__ __forceinline__ double dpp_row_shift_d(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, kCtrl, 0xF, 0xF, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), kCtrl, 0xF,
                                                              0xF, true);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double lane_value_d(double v, int l)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint64_t r = ((uint64_t)lane_value((uint32_t)(b >> 32), l) << 32) |
                       lane_value((uint32_t)b, l);
    return __longlong_as_double((long long)r);
}

__device__ __forceinline__ double row_scan_d(double v, int lane)
{
    v = v + dpp_row_shift_d<0x111>(v);    // row_shr:1 (0 shifted in)
    v = v + dpp_row_shift_d<0x112>(v);    // row_shr:2
    v = v + dpp_row_shift_d<0x114>(v);    // row_shr:4
    v = v + dpp_row_shift_d<0x118>(v);    // row_shr:8
    const double t0 = lane_value_d(v, 15), t1 = lane_value_d(v, 31), t2 = lane_value_d(v, 47);
    const double o2 = t0 + t1, o3 = o2 + t2;
    const int row = lane >> 4;
    if (row > 0) v = v + (row == 1 ? t0 : row == 2 ? o2 : o3);
    return v;
}

// The value row_scan_d leaves in lane 63 (the total of the 64 lanes), in every
// lane, without the scan's cross-row pass: the same four in-row steps, then
// ((t0 + t1) + t2) + t3 from the row totals (lane 63 holds t3 + ((t0 + t1) + t2),
// the same sum), read to scalar registers, so no LDS permute returns it.
__device__ __forceinline__ double row_total_d(double v)
{
    v = v + dpp_row_shift_d<0x111>(v);
    v = v + dpp_row_shift_d<0x112>(v);
    v = v + dpp_row_shift_d<0x114>(v);
    v = v + dpp_row_shift_d<0x118>(v);
    const double t0 = lane_value_d(v, 15), t1 = lane_value_d(v, 31), t2 = lane_value_d(v, 47),
                 t3 = lane_value_d(v, 63);
    return ((t0 + t1) + t2) + t3;
}

// w = exp(-alpha r) of a squared distance (every path computes it this one way).
__device__ __forceinline__ float w_of(uint32_t d2bits, float alpha_f)
{
    return exp_own(-(alpha_f * __builtin_sqrtf(__uint_as_float(d2bits))));
}

// What a cluster keeps in an electron's w slot: w itself, or (fixed) the fire
// threshold of x = fma(bdt, w, xA) as its 32-bit pattern, 0xFFFFFFFF for "always"
// (no other threshold reaches it: p < 1 in float32 gives at most 2^32 - 256), for
// runs in which that threshold is constant while w is (isothermal, one
// localisation channel).  The step then reads the threshold instead of taking
// exp of every electron's x again.
struct WSlot {
    bool fixed;
    float bdt, xA;
    __device__ __forceinline__ float of(float w) const
    {
        if (!fixed) return w;
        const FireThr t = fire_threshold(__builtin_fmaf(bdt, w, xA));
        return __uint_as_float(t.always ? 0xFFFFFFFFu : t.thr);
    }
};

__device__ __forceinline__ FireThr slot_threshold(float v)
{
    const uint32_t b = __float_as_uint(v);
    return FireThr{b, b == 0xFFFFFFFFu};
}

__device__ __forceinline__ void nn_from_key(uint64_t key, float alpha_f, uint32_t& slot, float& w)
{
    const uint32_t d2bits = (uint32_t)(key >> 32);
    const bool none = d2bits >= 0x7f800000u;        // no live hole left
    slot = none ? kNoHole : (uint32_t)(key & 0xFFFFu);
    w = none ? 0.0f : w_of(d2bits, alpha_f);
}

// Orders the LDS accesses of one wave around it, for a workgroup of one wave: a
// wavefront-scope fence for the compiler; the wave's LDS instructions execute in
// issue order.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bit r set iff any lane has bit r of its row mask set (rows <= 64), by ballots.
__device__ __forceinline__ uint64_t rows_any(uint64_t bits, int rows)
{
    uint64_t m = 0;
    for (int r = 0; r < rows; ++r)
        if (ballot((bits >> r) & 1ull) != 0) m |= 1ull << r;
    return m;
}

// Holes of one cluster counting-sorted by grid cell (Box.seed, engine.py:124-129)
// and every electron's initial nearest hole (Box._rebuild, engine.py:113-119):
// fills hole, cstart, est and ew.  NW wavefronts per workgroup (tid = thread
// index): the hole counts and the electrons' searches are split over the waves;
// the cell prefix, the scatter (slots within a cell in hole-index order) and the
// shift are wave 0's, between the workgroup's barriers.  Shared by the fixed-dt
// and event kernels.
template <int NW>
__device__ __forceinline__ Grid setup_cluster_t(const RunDesc& R, const ElectronGen& egen,
                                                float4* hole, uint32_t* cstart, uint32_t* est,
                                                float* ew, uint32_t prun, uint32_t psim,
                                                uint32_t k0, uint32_t k1, int tid,
                                                uint64_t& dist_evals,
                                                const WSlot& ws = WSlot{false, 0.0f, 0.0f})
{
    constexpr int NT = NW * kWave;
    const int lane = NW == 1 ? tid : (tid & (kWave - 1));
    const bool w0 = NW == 1 || tid < kWave;
    const int wv = NW == 1 ? 0 : uniform(tid >> 6);
    const int n_e0 = uniform(R.n_e0);
    const int n_h = uniform(R.n_h);
    const int G = uniform(R.grid);
    const int n_cells = G * G * G;
    const float cell = R.cell_nm, inv_cell = R.inv_cell;
    // ---- holes (Box.seed, engine.py:124-129), counting-sorted by grid cell ------
    for (int i = tid; i <= n_cells; i += NT) cstart[i] = 0u;
    __syncthreads();
    for (int h = tid; h < n_h; h += NT) {
        const u32x4 o = philox10(0u, (uint32_t)h, prun, psim | MCL_TAG_GEOM_H, k0, k1);
        const float x = unit_coord(o.x, R.side_b_nm), y = unit_coord(o.y, R.side_b_nm),
                    z = unit_coord(o.z, R.side_b_nm);
        const int cc = (cell_of(z, inv_cell, G) * G + cell_of(y, inv_cell, G)) * G +
                       cell_of(x, inv_cell, G);
        atomicAdd(&cstart[cc + 1], 1u);
    }
    __syncthreads();
    if (w0) {
        uint32_t carry = 0;
        for (int base = 0; base <= n_cells; base += kWave) {
            const int i = base + lane;
            uint32_t v = i <= n_cells ? cstart[i] : 0u;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const uint32_t u = (uint32_t)__shfl_up((int)v, off);
                if (lane >= off) v += u;
            }
            v += carry;
            if (i <= n_cells) cstart[i] = v;
            carry = (uint32_t)__shfl((int)v, kWave - 1);
        }
    }
    __syncthreads();
    if (w0) {
        for (int h = lane; h < n_h; h += kWave) {
            const u32x4 o = philox10(0u, (uint32_t)h, prun, psim | MCL_TAG_GEOM_H, k0, k1);
            const float x = unit_coord(o.x, R.side_b_nm), y = unit_coord(o.y, R.side_b_nm),
                        z = unit_coord(o.z, R.side_b_nm);
            const int cc = (cell_of(z, inv_cell, G) * G + cell_of(y, inv_cell, G)) * G +
                           cell_of(x, inv_cell, G);
            const uint32_t slot = atomicAdd(&cstart[cc], 1u);
            hole[slot] = make_float4(x, y, z, __uint_as_float(0xFFFF0000u | (uint32_t)h));
        }
    }
    __syncthreads();
    // cstart[c] now holds the end of cell c: shift right by one (top chunk first)
    for (int base = (n_cells / kWave) * kWave; base >= 0; base -= kWave) {
        const int i = base + lane;
        const uint32_t v = (w0 && i >= 1 && i <= n_cells) ? cstart[i - 1] : 0u;
        __syncthreads();
        if (w0 && i <= n_cells) cstart[i] = v;
        __syncthreads();
    }

    const Grid grid{hole, cstart, G, cell, inv_cell, n_h};
    const float alpha_f = (float)R.alpha_nm;

    // ---- electrons and their initial nearest hole (Box._rebuild, engine.py:113-119)
    const int rows0 = (n_e0 + kWave - 1) / kWave;
    for (int r = wv; r < rows0; r += NW) {
        const int e = r * kWave + lane;
        const bool act = e < n_e0;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (act) egen.xyz((uint32_t)e, x, y, z);
        const uint64_t key = nn_resolve(grid, act, x, y, z, lane, dist_evals);
        if (act) {
            uint32_t slot;
            float w;
            nn_from_key(key, alpha_f, slot, w);
            est[e] = (slot << 16) | (uint32_t)e;
            ew[e] = ws.of(w);
        }
    }
    __syncthreads();

    return grid;
}

__device__ __forceinline__ Grid setup_cluster(const RunDesc& R, const ElectronGen& egen,
                                              float4* hole, uint32_t* cstart, uint32_t* est,
                                              float* ew, uint32_t prun, uint32_t psim,
                                              uint32_t k0, uint32_t k1, int lane,
                                              uint64_t& dist_evals)
{
    return setup_cluster_t<1>(R, egen, hole, cstart, est, ew, prun, psim, k0, k1, lane, dist_evals);
}

// New nearest holes for the queued electrons rq[0..nq) ((slot << 16) | index):
// wave-cooperative searches, kSearchBatch electrons at a time, while the queue
// is short (at most wave_search_max), lane-parallel otherwise; writes est and ew
// of their slots (ew as ws keeps it).
__device__ __forceinline__ void search_queue_nn(const Grid& grid, const ElectronGen& egen,
                                                uint32_t* est, float* ew, const uint32_t* rq,
                                                int nq, int wave_search_max, float alpha_f,
                                                int lane, uint64_t& dist_evals,
                                                const WSlot& ws = WSlot{false, 0.0f, 0.0f})
{
    if (nq <= wave_search_max) {
        for (int i0 = 0; i0 < nq; i0 += kSearchBatch) {
            const int nb = min(kSearchBatch, nq - i0);
            if (nb == 1) {
                // one electron: the whole wave takes its block (9 ranges x 7 lanes)
                const uint32_t qq = rq[i0];
                float qx, qy, qz;
                egen.xyz(qq & 0xFFFFu, qx, qy, qz);
                bool ok = true;
                uint64_t key = block_search_wave(grid, qx, qy, qz, lane, ok, dist_evals);
                if (!ok) key = full_scan_wave(grid, qx, qy, qz, lane, dist_evals);
                if (lane == 0) {
                    uint32_t slot;
                    float w;
                    nn_from_key(key, alpha_f, slot, w);
                    est[qq >> 16] = (slot << 16) | (qq & 0xFFFFu);
                    ew[qq >> 16] = ws.of(w);
                }
                continue;
            }
            // two to four electrons: one 16-lane row each; every lane of row b
            // regenerates electron b's coordinates (the same Philox call in all lanes)
            const int b = lane >> 4;
            const bool act = b < nb;
            const uint32_t qq = act ? rq[i0 + b] : 0u;
            float qx = 0.0f, qy = 0.0f, qz = 0.0f;
            if (act) egen.xyz(qq & 0xFFFFu, qx, qy, qz);
            uint64_t key = block_search_rows(grid, nb, qx, qy, qz, lane, dist_evals);
            bool ok = true;
            if (act) {
                const Block k(grid, qx, qy, qz);
                ok = k.conclusive(grid, qx, qy, qz, __uint_as_float((uint32_t)(key >> 32)));
            }
            // rows whose block was inconclusive: the whole wave scans every hole
            uint64_t fb = ballot((lane & 15) == 0 && act && !ok);
            while (fb != 0) {
                const int l = __builtin_ctzll(fb);
                fb &= fb - 1;
                const uint64_t k = full_scan_wave(grid, lane_value(qx, l), lane_value(qy, l),
                                                  lane_value(qz, l), lane, dist_evals);
                if ((lane >> 4) == (l >> 4)) key = k;
            }
            if (act && (lane & 15) == 0) {
                uint32_t slot;
                float w;
                nn_from_key(key, alpha_f, slot, w);
                est[qq >> 16] = (slot << 16) | (qq & 0xFFFFu);
                ew[qq >> 16] = ws.of(w);
            }
        }
        return;
    }
    for (int q0 = 0; q0 < nq; q0 += kWave) {
        const int i = q0 + lane;
        const bool act = i < nq;
        const uint32_t q = act ? rq[i] : 0u;
        const uint32_t o = q & 0xFFFFu;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (act) egen.xyz(o, x, y, z);
        const uint64_t key = nn_resolve(grid, act, x, y, z, lane, dist_evals);
        if (act) {
            uint32_t slot;
            float w;
            nn_from_key(key, alpha_f, slot, w);
            est[q >> 16] = (slot << 16) | o;
            ew[q >> 16] = ws.of(w);
        }
    }
}

}  // namespace mcl
