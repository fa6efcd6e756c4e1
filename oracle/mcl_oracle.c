/*
 * mcl_oracle.c — CPU replay harness for the fixed-dt trap-kinetics Monte Carlo.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py may load liborc.so; the product
 * (libmcl.so and the mcluminescence_amd package) never does.
 *
 * It restates, in plain C and independently of the HIP kernels, the algorithm
 * of the reference's hot path with the fixed-dt step the kernels implement:
 *   geometry       tl_trap_lab.py:33-39, engine.py:124-129
 *   nearest hole   engine.py:113-119 (argmin, first index on ties), 154-175
 *   rates          engine.py:65-77  (k = s e^{-E_cb/kT} + b e^{-E_loc/kT - alpha r},
 *                                    E_loc = E_loc_2 if U < Retrap else E_loc_1)
 *   step loop      simulate.py:46-92 (start-of-step temperature)
 *   conflicts      src/est_params/functions.py:228-239 (first electron keeps the hole);
 *                  a loser re-searches its nearest hole and retries within the same
 *                  step, accepted with probability T_new / T_old (thinning), up to
 *                  MCL_MAX_ROUNDS resolution rounds per step
 * Random numbers come from Philox4x32-10 (Salmon et al., SC'11) keyed by the
 * run seed and counted by (step block, particle, run, sim|tag); transcendental
 * functions are this file's own Cody-Waite + Taylor evaluations with explicit
 * fma, so that every float32 decision is reproducible bit for bit.
 *
 * Mode fp64 = 1 keeps coordinates, distances and probabilities in float64 with
 * libm exp/expm1: the all-FP64 reference against which the float32 curves'
 * error is stated.  It consumes the same Philox stream.
 *
 * Parity pins (tests/): Philox against Random123 known answers and rocRAND's
 * independent implementation; exp/expm1 against libm to <= 2 ulp; geometry
 * against the closed forms; nearest-hole maintenance against brute force;
 * ensembles against closed forms and the numpy event-mode restatement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/mcl.h"

/* ---------------------------------------------------------------- Philox -- */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += PH_W0; k1 += PH_W1; }
        uint64_t p0 = (uint64_t)PH_M0 * x0;
        uint64_t p1 = (uint64_t)PH_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

static inline uint32_t philox_word(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                   const uint32_t key[2], int word)
{
    uint32_t ctr[4] = {c0, c1, c2, c3}, out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[word];
}

/* ------------------------------------------------------- own exp / expm1 -- */
/* float32: x = k ln2 + r, |r| <= ln2/2, ln2 split in 12 + 24 bits (k*hi exact),
 * expm1(r) = r + r^2 (1/2! + r/3! + ... + r^6/8!) by Horner with fma.          */
#define F_INV_LN2 0x1.715476p+0f
#define F_LN2_HI  0x1.62ep-1f
#define F_LN2_LO  0x1.0bfbe8p-15f
#define F_C2 0x1p-1f
#define F_C3 0x1.555556p-3f
#define F_C4 0x1.555556p-5f
#define F_C5 0x1.111112p-7f
#define F_C6 0x1.6c16c2p-10f
#define F_C7 0x1.a01a02p-13f
#define F_C8 0x1.a01a02p-16f

static inline float f_pow2i(int k)   /* 2^k, -126 <= k <= 127 */
{
    uint32_t bits = (uint32_t)(k + 127) << 23;
    float f;
    memcpy(&f, &bits, sizeof f);
    return f;
}

static inline float f_em1_core(float r)
{
    float p = F_C8;
    p = fmaf(p, r, F_C7);
    p = fmaf(p, r, F_C6);
    p = fmaf(p, r, F_C5);
    p = fmaf(p, r, F_C4);
    p = fmaf(p, r, F_C3);
    p = fmaf(p, r, F_C2);
    float r2 = r * r;
    return fmaf(r2, p, r);
}

float orc_expm1f(float x)
{
    if (x != x) return x;
    if (x < -17.5f) return -1.0f;             /* e^x < 2^-25: rounds to -1 */
    if (x > 88.0f) return INFINITY;
    float kf = rintf(x * F_INV_LN2);
    float r = fmaf(-kf, F_LN2_HI, x);
    r = fmaf(-kf, F_LN2_LO, r);
    float em1 = f_em1_core(r);
    if (kf == 0.0f) return em1;
    float t = f_pow2i((int)kf);
    return fmaf(t, em1, t - 1.0f);
}

float orc_expf(float x)
{
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;              /* results below FLT_MIN flush to 0 */
    if (x > 88.0f) return INFINITY;
    float kf = rintf(x * F_INV_LN2);
    float r = fmaf(-kf, F_LN2_HI, x);
    r = fmaf(-kf, F_LN2_LO, r);
    float e = 1.0f + f_em1_core(r);
    return e * f_pow2i((int)kf);
}

/* float64: same reduction (ln2 split 21 + 53 bits), exp(r) Taylor to r^13/13!. */
#define D_INV_LN2 0x1.71547652b82fep+0
#define D_LN2_HI  0x1.62e42p-1
#define D_LN2_LO  0x1.fdf473de6af28p-22

static const double D_INVFACT[14] = {
    0x1p+0, 0x1p+0, 0x1p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
    0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
    0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
    0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33,
};

static inline double d_pow2i(int k)  /* 2^k, -1022 <= k <= 1023 */
{
    uint64_t bits = (uint64_t)(k + 1023) << 52;
    double d;
    memcpy(&d, &bits, sizeof d);
    return d;
}

double orc_exp(double x)
{
    if (x != x) return x;
    if (x < -708.0) return 0.0;               /* results below DBL_MIN flush to 0 */
    if (x > 709.0) return INFINITY;
    double kd = rint(x * D_INV_LN2);
    double r = fma(-kd, D_LN2_HI, x);
    r = fma(-kd, D_LN2_LO, r);
    double p = D_INVFACT[13];
    for (int i = 12; i >= 1; --i) p = fma(p, r, D_INVFACT[i]);
    double e = fma(p, r, 1.0);
    return e * d_pow2i((int)kd);
}

/* float64 natural log for normal x > 0: x = 2^k m with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m - 1) / (m + 1), by the odd series 2 s (1 + s^2/3 + ...
 * + s^22/23) (|s| <= 0.1716), then k ln 2 with the same two-part ln 2 as exp. */
double orc_log(double x)
{
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    int k = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &b, sizeof m);
    if (m > 0x1.6a09e667f3bcdp+0) { m *= 0.5; k += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double q = s * s;
    double P = 1.0 / 23.0;
    P = fma(P, q, 1.0 / 21.0);
    P = fma(P, q, 1.0 / 19.0);
    P = fma(P, q, 1.0 / 17.0);
    P = fma(P, q, 1.0 / 15.0);
    P = fma(P, q, 1.0 / 13.0);
    P = fma(P, q, 1.0 / 11.0);
    P = fma(P, q, 1.0 / 9.0);
    P = fma(P, q, 1.0 / 7.0);
    P = fma(P, q, 1.0 / 5.0);
    P = fma(P, q, 1.0 / 3.0);
    double t = 2.0 * s;
    double lnm = fma(t, q * P, t);
    double kd = (double)k;
    return fma(kd, D_LN2_HI, fma(kd, D_LN2_LO, lnm));
}

/* -------------------------------------------------------------- geometry -- */
int orc_run_geometry(const mcl_run_params* p, mcl_geometry* g)
{
    /* tl_trap_lab.py:33-34: rho = rho' * (3/(4 pi) * alpha^3); d = (holes/rho)^(1/3) */
    double rho = p->rho_prime * (3.0 / (4.0 * 0x1.921fb54442d18p+1) * pow(p->alpha, 3.0));
    g->rho = rho;
    g->side_m = pow(p->holes / rho, 1.0 / 3.0);
    /* tl_trap_lab.py:38, engine.py:127: int() truncation */
    g->n_e0 = (int32_t)(p->N_e * p->e_ratio_start);
    g->n_h = (int32_t)(p->holes * pow(p->boundary_factor, 3.0));
    return 0;
}

/* initialize_box_bg (functions.py:51-80): side from int(holes), the boundary shell
 * holds int(int(holes) / d^3 * ((bf d)^3 - d^3)) more holes. */
int orc_run_geometry_legacy(const mcl_run_params* p, mcl_geometry* g)
{
    double rho = p->rho_prime * (3.0 / (4.0 * 0x1.921fb54442d18p+1) * pow(p->alpha, 3.0));
    double h = trunc(p->holes);
    double d = pow(h / rho, 1.0 / 3.0);
    double bl = d * p->boundary_factor;
    double box = d * d * d;
    double box_b = bl * bl * bl;
    g->rho = rho;
    g->side_m = d;
    g->n_e0 = (int32_t)(p->N_e * p->e_ratio_start);
    g->n_h = (int32_t)h + (int32_t)((h / (d * d * d)) * (box_b - box));
    return 0;
}

/* -------------------------------------------------------- cluster state --- */
typedef struct {
    int n_e0, n_h;
    float *ex, *ey, *ez, *hx, *hy, *hz;        /* float32 mode */
    double *dex, *dey, *dez, *dhx, *dhy, *dhz; /* float64 mode */
    unsigned char* h_alive;
    uint32_t* nn;
    float* w;
    double* dw;
    int32_t* live;    /* live electron indices, ascending */
    int n_live;
} cluster_t;

static void* xcalloc(size_t n, size_t sz) { void* p = calloc(n ? n : 1, sz); return p; }

static void cluster_free(cluster_t* c)
{
    free(c->ex); free(c->ey); free(c->ez); free(c->hx); free(c->hy); free(c->hz);
    free(c->dex); free(c->dey); free(c->dez); free(c->dhx); free(c->dhy); free(c->dhz);
    free(c->h_alive); free(c->nn); free(c->w); free(c->dw); free(c->live);
}

/* nearest live hole of electron e: float32 squared distance
 * d2 = fma(dz, dz, fma(dy, dy, dx*dx)); strict < keeps the first index on ties */
static uint32_t nearest_f32(const cluster_t* c, int e, float* d2_out)
{
    float best = INFINITY;
    uint32_t idx = MCL_NO_HOLE;
    const float x = c->ex[e], y = c->ey[e], z = c->ez[e];
    for (int h = 0; h < c->n_h; ++h) {
        if (!c->h_alive[h]) continue;
        float dx = c->hx[h] - x, dy = c->hy[h] - y, dz = c->hz[h] - z;
        float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (d2 < best) { best = d2; idx = (uint32_t)h; }
    }
    *d2_out = best;
    return idx;
}

static uint32_t nearest_f64(const cluster_t* c, int e, double* d2_out)
{
    double best = INFINITY;
    uint32_t idx = MCL_NO_HOLE;
    const double x = c->dex[e], y = c->dey[e], z = c->dez[e];
    for (int h = 0; h < c->n_h; ++h) {
        if (!c->h_alive[h]) continue;
        double dx = c->dhx[h] - x, dy = c->dhy[h] - y, dz = c->dhz[h] - z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; idx = (uint32_t)h; }
    }
    *d2_out = best;
    return idx;
}

static void refresh_nn(cluster_t* c, int e, float alpha_f, double alpha_nm, int fp64)
{
    if (fp64) {
        double d2;
        c->nn[e] = nearest_f64(c, e, &d2);
        c->dw[e] = (c->nn[e] == MCL_NO_HOLE) ? 0.0 : exp(-(alpha_nm * sqrt(d2)));
    } else {
        float d2;
        c->nn[e] = nearest_f32(c, e, &d2);
        c->w[e] = (c->nn[e] == MCL_NO_HOLE) ? 0.0f : orc_expf(-(alpha_f * sqrtf(d2)));
    }
}

static int cluster_init(cluster_t* c, const mcl_run_params* p, int run, int sim,
                        const uint32_t key[2], int fp64)
{
    mcl_geometry g;
    orc_run_geometry(p, &g);
    memset(c, 0, sizeof *c);
    c->n_e0 = g.n_e0 > 0 ? g.n_e0 : 0;
    c->n_h = g.n_h > 0 ? g.n_h : 0;
    const double side_nm = g.side_m * 1e9;
    const double side_b_nm = (g.side_m * p->boundary_factor) * 1e9;
    const uint32_t c3e = ((uint32_t)sim << 4) | MCL_TAG_GEOM_E;
    const uint32_t c3h = ((uint32_t)sim << 4) | MCL_TAG_GEOM_H;
    int ne = c->n_e0, nh = c->n_h;
    c->h_alive = xcalloc(nh, 1);
    c->nn = xcalloc(ne, sizeof(uint32_t));
    c->live = xcalloc(ne, sizeof(int32_t));
    if (fp64) {
        c->dex = xcalloc(ne, 8); c->dey = xcalloc(ne, 8); c->dez = xcalloc(ne, 8);
        c->dhx = xcalloc(nh, 8); c->dhy = xcalloc(nh, 8); c->dhz = xcalloc(nh, 8);
        c->dw = xcalloc(ne, 8);
    } else {
        c->ex = xcalloc(ne, 4); c->ey = xcalloc(ne, 4); c->ez = xcalloc(ne, 4);
        c->hx = xcalloc(nh, 4); c->hy = xcalloc(nh, 4); c->hz = xcalloc(nh, 4);
        c->w = xcalloc(ne, 4);
    }
    for (int e = 0; e < ne; ++e) {
        uint32_t ctr[4] = {0u, (uint32_t)e, (uint32_t)run, c3e}, o[4];
        orc_philox4x32_10(ctr, key, o);
        double ux = (double)o[0] * 0x1p-32, uy = (double)o[1] * 0x1p-32, uz = (double)o[2] * 0x1p-32;
        if (fp64) { c->dex[e] = ux * side_nm; c->dey[e] = uy * side_nm; c->dez[e] = uz * side_nm; }
        else { c->ex[e] = (float)(ux * side_nm); c->ey[e] = (float)(uy * side_nm); c->ez[e] = (float)(uz * side_nm); }
        c->live[e] = e;
    }
    for (int h = 0; h < nh; ++h) {
        uint32_t ctr[4] = {0u, (uint32_t)h, (uint32_t)run, c3h}, o[4];
        orc_philox4x32_10(ctr, key, o);
        double ux = (double)o[0] * 0x1p-32, uy = (double)o[1] * 0x1p-32, uz = (double)o[2] * 0x1p-32;
        if (fp64) { c->dhx[h] = ux * side_b_nm; c->dhy[h] = uy * side_b_nm; c->dhz[h] = uz * side_b_nm; }
        else { c->hx[h] = (float)(ux * side_b_nm); c->hy[h] = (float)(uy * side_b_nm); c->hz[h] = (float)(uz * side_b_nm); }
        c->h_alive[h] = 1;
    }
    c->n_live = ne;
    const double alpha_nm = p->alpha * 1e-9;
    const float alpha_f = (float)alpha_nm;
    for (int e = 0; e < ne; ++e) refresh_nn(c, e, alpha_f, alpha_nm, fp64);
    return 0;
}

/* ------------------------------------------------------------ step rule --- */
/* Fire threshold T in [0, 2^32] for x = k dt: p = 1 - e^{-x}; T = 2^32 when
 * p >= 1, else (uint32)(p * 2^32).  An electron with draw u fires iff u < T. */
static inline uint64_t thr_f32(float x)
{
    float p = -orc_expm1f(-x);
    return p >= 1.0f ? 0x100000000ull : (uint64_t)(uint32_t)(p * 0x1p32f);
}

static inline uint64_t thr_f64(double x)
{
    double p = -expm1(-x);
    return p >= 1.0 ? 0x100000000ull : (uint64_t)(p * 0x1p32);
}

typedef struct {
    double A, B1, B2;         /* s e^{-E_cb/kT}, b e^{-E_loc_i/kT}           */
    float xA, b1dt, b2dt;     /* float32 A dt, B_i dt                          */
} step_scalars_t;

static inline uint64_t electron_thr(const cluster_t* c, const mcl_run_params* p,
                                    const step_scalars_t* s, int e, int chan2, int fp64)
{
    if (fp64) return thr_f64(s->A * p->dt + (chan2 ? s->B2 : s->B1) * p->dt * c->dw[e]);
    return thr_f32(fmaf(chan2 ? s->b2dt : s->b1dt, c->w[e], s->xA));
}

static int validate(const mcl_run_params* p, int n_bins)
{
    if (!(p->dt > 0) || p->n_steps <= 0 || p->sims < 0) return MCL_EINVAL;
    if (n_bins < 0) return MCL_EINVAL;
    if (n_bins > p->n_steps) return MCL_EOVERFLOW;
    if (p->D != 0.0) return MCL_EINVAL;
    if (!(p->k_b > 0) || !(p->T_start + 273.15 > 0)) return MCL_EINVAL;
    mcl_geometry g;
    orc_run_geometry(p, &g);
    if (g.n_e0 > 65535 || g.n_h > 65534) return MCL_ERANGE;
    return 0;
}

/* Final state of a cluster after a bounded number of steps (test hook). */
typedef struct {
    int max_steps;            /* stop after this many steps                     */
    unsigned char* e_live;    /* [n_e0] 1 if the electron is still trapped      */
    unsigned char* h_alive;   /* [n_h]                                          */
    uint32_t* nn;             /* [n_e0] maintained nearest hole                 */
    float* e_xyz;             /* [n_e0][3]                                      */
    float* h_xyz;             /* [n_h][3]                                       */
} final_state_t;

/* Simulate one cluster; counts[row_len] (zeroed here), returns live electron-steps.
 * n_bins == 0: one bin per step. */
static uint64_t run_cluster(const mcl_run_params* p, int run, int sim, const uint32_t key[2],
                            int n_bins, int row_len, int fp64, uint32_t* counts, final_state_t* fs)
{
    cluster_t c;
    cluster_init(&c, p, run, sim, key, fp64);
    memset(counts, 0, sizeof(uint32_t) * (size_t)row_len);
    const int n_steps = p->n_steps;
    if (n_bins == 0) n_bins = n_steps;
    const double T0_K = p->T_start + 273.15;
    const double alpha_nm = p->alpha * 1e-9;
    const float alpha_f = (float)alpha_nm;
    const int two_channel = (p->E_loc_1 != p->E_loc_2) && (p->Retrap > 0.0) && (p->Retrap < 1.0);
    const int always_chan2 = p->Retrap >= 1.0;
    const uint64_t ret_thr = two_channel ? (uint64_t)(p->Retrap * 0x1p32) : 0;
    const uint32_t c3f = ((uint32_t)sim << 4) | MCL_TAG_FIRE;
    const uint32_t c3c = ((uint32_t)sim << 4) | MCL_TAG_CHAN;

    int32_t* cand = xcalloc(c.n_e0, sizeof(int32_t));
    int32_t* lose = xcalloc(c.n_e0, sizeof(int32_t));
    uint64_t* thr_prev = xcalloc(c.n_e0, sizeof(uint64_t));
    uint32_t* claim = xcalloc(c.n_h, sizeof(uint32_t));
    for (int h = 0; h < c.n_h; ++h) claim[h] = 0xFFFFFFFFu;
    unsigned char* won = xcalloc(c.n_e0, 1);
    uint64_t live_steps = 0;

    const int step_limit = (fs && fs->max_steps < n_steps) ? fs->max_steps : n_steps;
    for (int n = 0; n < step_limit && c.n_live > 0; ++n) {
        live_steps += (uint64_t)c.n_live;
        /* per-step float64 scalars (simulate.py:53-54, engine.py:65-73) */
        const double t = (double)n * p->dt;
        const double T = T0_K + p->T_rate * t;
        const double kT = p->k_b * T;
        double A, B1, B2;
        if (fp64) {
            A = p->s * exp(-(p->E_cb / kT));
            B1 = p->b * exp(-(p->E_loc_1 / kT));
            B2 = p->b * exp(-(p->E_loc_2 / kT));
        } else {
            A = p->s * orc_exp(-(p->E_cb / kT));
            B1 = p->b * orc_exp(-(p->E_loc_1 / kT));
            B2 = p->b * orc_exp(-(p->E_loc_2 / kT));
        }
        const uint32_t blk = (uint32_t)n >> 2;
        const int word = n & 3;
        step_scalars_t sc = {A, B1, B2, (float)(A * p->dt), (float)(B1 * p->dt),
                             (float)(B2 * p->dt)};
        /* round 0 candidates: electrons whose draw falls below their threshold */
        int n_cand = 0;
        for (int i = 0; i < c.n_live; ++i) {
            const int e = c.live[i];
            if (c.nn[e] == MCL_NO_HOLE) continue;
            int chan2 = always_chan2;
            if (two_channel) {
                uint32_t uc = philox_word(blk, (uint32_t)e, (uint32_t)run, c3c, key, word);
                chan2 = (uint64_t)uc < ret_thr;
            }
            const uint32_t u = philox_word(blk, (uint32_t)e, (uint32_t)run, c3f, key, word);
            const uint64_t T = electron_thr(&c, p, &sc, e, chan2, fp64);
            if ((uint64_t)u < T) { cand[n_cand++] = e; thr_prev[e] = T; }   /* ascending */
        }
        if (n_cand == 0) continue;
        int wins = 0;
        for (int round = 0; round < MCL_MAX_ROUNDS && n_cand > 0; ++round) {
            /* the lowest electron index keeps a shared hole (functions.py:228-239) */
            int n_lose = 0;
            for (int j = 0; j < n_cand; ++j) {
                const int e = cand[j];
                const uint32_t h = c.nn[e];
                if (claim[h] == 0xFFFFFFFFu) { claim[h] = (uint32_t)e; won[e] = 1; ++wins; }
                else lose[n_lose++] = e;
            }
            for (int j = 0; j < n_cand; ++j) {
                const int e = cand[j];
                if (won[e]) c.h_alive[c.nn[e]] = 0;
                claim[c.nn[e]] = 0xFFFFFFFFu;
            }
            /* remove winners (stable), re-search electrons whose hole died */
            int k = 0;
            for (int i = 0; i < c.n_live; ++i) {
                const int e = c.live[i];
                if (won[e]) { won[e] = 0; continue; }
                c.live[k++] = e;
            }
            c.n_live = k;
            for (int i = 0; i < c.n_live; ++i) {
                const int e = c.live[i];
                if (c.nn[e] != MCL_NO_HOLE && !c.h_alive[c.nn[e]])
                    refresh_nn(&c, e, alpha_f, alpha_nm, fp64);
            }
            /* a loser retries on its new nearest hole in the same step, accepted
             * with probability T_new / T_old (thinning of the draw it already won) */
            n_cand = 0;
            if (round + 1 == MCL_MAX_ROUNDS) break;
            const uint32_t c3r = ((uint32_t)sim << 4) | (MCL_TAG_RETRY0 + (uint32_t)round);
            for (int j = 0; j < n_lose; ++j) {
                const int e = lose[j];
                if (c.nn[e] == MCL_NO_HOLE) continue;
                int chan2 = always_chan2;
                if (two_channel) {
                    uint32_t uc = philox_word(blk, (uint32_t)e, (uint32_t)run, c3c, key, word);
                    chan2 = (uint64_t)uc < ret_thr;
                }
                const uint64_t Tn = electron_thr(&c, p, &sc, e, chan2, fp64);
                const uint64_t To = thr_prev[e];
                int accept = Tn >= To;
                if (!accept) {
                    const uint32_t u2 = philox_word(blk, (uint32_t)e, (uint32_t)run, c3r, key, word);
                    accept = (uint64_t)u2 * To < (Tn << 32);
                }
                if (accept) { cand[n_cand++] = e; thr_prev[e] = Tn; }
            }
        }
        const int bin = (int)(((int64_t)n * n_bins) / n_steps);
        counts[bin] += (uint32_t)wins;
    }
    if (fs) {
        for (int e = 0; e < c.n_e0; ++e) fs->e_live[e] = 0;
        for (int i = 0; i < c.n_live; ++i) fs->e_live[c.live[i]] = 1;
        for (int h = 0; h < c.n_h; ++h) fs->h_alive[h] = c.h_alive[h];
        for (int e = 0; e < c.n_e0; ++e) fs->nn[e] = c.nn[e];
        if (!fp64) {
            for (int e = 0; e < c.n_e0; ++e) {
                fs->e_xyz[3 * e] = c.ex[e]; fs->e_xyz[3 * e + 1] = c.ey[e]; fs->e_xyz[3 * e + 2] = c.ez[e];
            }
            for (int h = 0; h < c.n_h; ++h) {
                fs->h_xyz[3 * h] = c.hx[h]; fs->h_xyz[3 * h + 1] = c.hy[h]; fs->h_xyz[3 * h + 2] = c.hz[h];
            }
        }
    }
    free(cand); free(lose); free(thr_prev); free(claim); free(won);
    cluster_free(&c);
    return live_steps;
}

/* counts [total_clusters][n_bins]; live_steps [total_clusters] or NULL. */
int orc_run_fixed_dt(const mcl_run_params* runs, int32_t n_runs, uint64_t seed,
                     int32_t n_bins, int32_t fp64, int32_t n_threads,
                     uint32_t* counts, uint64_t* live_steps)
{
    if (!runs || n_runs <= 0 || !counts) return MCL_EINVAL;
    int64_t total = 0;
    int row_len = n_bins;
    for (int r = 0; r < n_runs; ++r) {
        int st = validate(&runs[r], n_bins);
        if (st) return st;
        total += runs[r].sims;
        if (n_bins == 0 && runs[r].sims > 0 && runs[r].n_steps > row_len) row_len = runs[r].n_steps;
    }
    int32_t* c_run = xcalloc((size_t)total, sizeof(int32_t));
    int32_t* c_sim = xcalloc((size_t)total, sizeof(int32_t));
    int64_t k = 0;
    for (int r = 0; r < n_runs; ++r)
        for (int s = 0; s < runs[r].sims; ++s) { c_run[k] = r; c_sim[k] = s; ++k; }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < total; ++i) {
        uint64_t ls = run_cluster(&runs[c_run[i]], c_run[i], c_sim[i], key, n_bins, row_len, fp64,
                                  counts + (size_t)i * (size_t)row_len, NULL);
        if (live_steps) live_steps[i] = ls;
    }
    free(c_run); free(c_sim);
    return 0;
}

/* Initial float32 state of one cluster (coordinates in nm, nn, w). */
int orc_init_state(const mcl_run_params* p, int32_t run, int32_t sim, uint64_t seed,
                   float* e_xyz, float* h_xyz, uint32_t* nn, float* w)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    cluster_t c;
    cluster_init(&c, p, run, sim, key, 0);
    for (int e = 0; e < c.n_e0; ++e) {
        e_xyz[3 * e] = c.ex[e]; e_xyz[3 * e + 1] = c.ey[e]; e_xyz[3 * e + 2] = c.ez[e];
        nn[e] = c.nn[e]; w[e] = c.w[e];
    }
    for (int h = 0; h < c.n_h; ++h) {
        h_xyz[3 * h] = c.hx[h]; h_xyz[3 * h + 1] = c.hy[h]; h_xyz[3 * h + 2] = c.hz[h];
    }
    cluster_free(&c);
    return 0;
}

/* State of cluster (run, sim) after max_steps float32 steps: which electrons
 * and holes are left and the incrementally maintained nearest hole. */
int orc_final_state(const mcl_run_params* p, int32_t run, int32_t sim, uint64_t seed,
                    int32_t max_steps, int32_t n_bins, uint32_t* counts,
                    unsigned char* e_live, unsigned char* h_alive, uint32_t* nn,
                    float* e_xyz, float* h_xyz)
{
    int st = validate(p, n_bins);
    if (st) return st;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    final_state_t fs = {max_steps, e_live, h_alive, nn, e_xyz, h_xyz};
    run_cluster(p, run, sim, key, n_bins, n_bins ? n_bins : p->n_steps, 0, counts, &fs);
    return 0;
}

/* Brute-force checker for nearest-hole maintenance: given coordinates and an
 * alive mask, recompute every electron's nearest hole from scratch. */
void orc_bruteforce_nn(const float* e_xyz, int n_e, const float* h_xyz, const unsigned char* alive,
                       int n_h, uint32_t* nn)
{
    for (int e = 0; e < n_e; ++e) {
        float best = INFINITY;
        uint32_t idx = MCL_NO_HOLE;
        for (int h = 0; h < n_h; ++h) {
            if (!alive[h]) continue;
            float dx = h_xyz[3 * h] - e_xyz[3 * e], dy = h_xyz[3 * h + 1] - e_xyz[3 * e + 1];
            float dz = h_xyz[3 * h + 2] - e_xyz[3 * e + 2];
            float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (d2 < best) { best = d2; idx = (uint32_t)h; }
        }
        nn[e] = idx;
    }
}

/* Monotonicity of the float32 fire threshold: walks every float x in [lo, hi]
 * in increasing order and counts the places where T(x) falls below T at a
 * smaller x (the running maximum).  *worst gets the largest such drop. */
uint64_t orc_thr_monotone_scan(float lo, float hi, uint64_t* worst)
{
    uint64_t bad = 0, run_max = 0;
    *worst = 0;
    uint32_t b = 0, e = 0;
    memcpy(&b, &lo, 4);
    memcpy(&e, &hi, 4);
    for (uint32_t i = b; i <= e; ++i) {
        float x;
        memcpy(&x, &i, 4);
        const uint64_t t = thr_f32(x);
        if (t < run_max) {
            ++bad;
            if (run_max - t > *worst) *worst = run_max - t;
        } else {
            run_max = t;
        }
        if (i == 0xFFFFFFFFu) break;
    }
    return bad;
}

/* The float32 fire threshold T(x) in [0, 2^32] (2^32: always fires). */
uint64_t orc_fire_threshold(float x) { return thr_f32(x); }

/* ------------------------------------------------------------ event mode -- */
/* Replay of the event kernel (mcluminescence_amd/csrc/mc_event.hip), the
 * reference's loop simulate.py:46-92 with one event per step:
 *   T = (T_start + T_rate t) + 273.15; A, B_c in float64 (own exp);
 *   k_e = fma(B_ch, w_e, A) per live electron in index order;
 *   K = rows of 64 electrons, each an inclusive scan (steps of 1, 2, 4, 8 within
 *   16-lane rows, then rows 1..3 add t0, t0 + t1, (t0 + t1) + t2 of the rows
 *   before), summed in order; one Philox block (i, 0xFFFFFFFF, run, sim|FIRE) gives
 *   E_t = -ln((u0 + 1/2) 2^-32), U_w = u1 2^-32, E_f = -ln((u2 + 1/2) 2^-32);
 *   dt_fill = (1 / lambda) E_f, dt_rec = E_t / K (dt_fill without electrons, +inf
 *   when K = 0), dt = min(dt_fill, dt_rec, dt_cap); a recombination takes the
 *   first electron whose running sum exceeds U_w K. */
static void scan64(double* v)
{
    double old[64];
    for (int o = 1; o < 16; o <<= 1) {
        memcpy(old, v, sizeof old);
        for (int l = 0; l < 64; ++l)
            if ((l & 15) >= o) v[l] = old[l] + old[l - o];
    }
    const double t0 = v[15], t1 = v[31], t2 = v[47];
    const double o2 = t0 + t1, o3 = o2 + t2;
    for (int l = 16; l < 64; ++l) v[l] = v[l] + (l < 32 ? t0 : l < 48 ? o2 : o3);
}

typedef struct {
    const cluster_t* c;
    const uint32_t* key;
    uint32_t run, c3c;
    int i, two_ch, always2;
    uint64_t ret_thr;
    double A, B1, B2;
} ev_rates_t;

static double ev_rate(const ev_rates_t* R, int slot)
{
    if (slot >= R->c->n_live) return 0.0;
    const int e = R->c->live[slot];
    if (R->c->nn[e] == MCL_NO_HOLE) return 0.0;
    int ch = R->always2;
    if (R->two_ch)
        ch = (uint64_t)philox_word((uint32_t)R->i >> 2, (uint32_t)e, R->run, R->c3c, R->key,
                                   R->i & 3) < R->ret_thr;
    return fma(ch ? R->B2 : R->B1, (double)R->c->w[e], R->A);
}

static void ev_row(const ev_rates_t* R, int r, double* v)
{
    for (int l = 0; l < 64; ++l) v[l] = ev_rate(R, r * 64 + l);
    scan64(v);
}

static int32_t run_cluster_event(const mcl_run_params* p, int run, int sim, const uint32_t key[2],
                                 double duration, int max_steps, uint8_t* lum, double* tax,
                                 double* erat, uint64_t* live_steps)
{
    cluster_t c;
    cluster_init(&c, p, run, sim, key, 0);
    const double alpha_nm = p->alpha * 1e-9;
    const float alpha_f = (float)alpha_nm;
    const int two_ch = (p->E_loc_1 != p->E_loc_2) && (p->Retrap > 0.0) && (p->Retrap < 1.0);
    ev_rates_t R = {&c, key, (uint32_t)run, ((uint32_t)sim << 4) | MCL_TAG_CHAN, 0, two_ch,
                    p->Retrap >= 1.0, two_ch ? (uint64_t)(p->Retrap * 0x1p32) : 0u, 0, 0, 0};
    const int b1_is_a = p->b == p->s && p->E_loc_1 == p->E_cb;
    const int b2_is_b1 = p->E_loc_2 == p->E_loc_1;
    const double dt_cap = p->T_rate != 0.0 ? p->dt / p->T_rate : 1e20;
    const int N_int = (int)p->N_e;
    const uint32_t c3f = ((uint32_t)sim << 4) | MCL_TAG_FIRE;
    double* rtot = xcalloc((size_t)(c.n_e0 / 64 + 2), sizeof(double));
    double v[64];
    int filled = 0, i = 0;
    int32_t status = 0;
    double t = 0.0;
    uint64_t ls = 0;
    memset(lum, 0, (size_t)max_steps);
    if (tax) memset(tax, 0, sizeof(double) * (size_t)max_steps);
    if (erat) memset(erat, 0, sizeof(double) * (size_t)max_steps);
    while (t <= duration) {
        if (i >= max_steps) { status = -1 - i; break; }
        ls += (uint64_t)c.n_live;
        const double T = (p->T_start + p->T_rate * t) + 273.15;
        const double kT = p->k_b * T;
        R.A = p->s * orc_exp(-(p->E_cb / kT));
        R.B1 = b1_is_a ? R.A : p->b * orc_exp(-(p->E_loc_1 / kT));
        R.B2 = b2_is_b1 ? R.B1 : p->b * orc_exp(-(p->E_loc_2 / kT));
        R.i = i;
        const int rows = (c.n_live + 63) / 64;
        double K = 0.0;
        for (int r = 0; r < rows; ++r) {
            ev_row(&R, r, v);
            rtot[r] = v[63];
            K = K + v[63];
        }
        uint32_t sb[4];
        {
            uint32_t ctr[4] = {(uint32_t)i, 0xFFFFFFFFu, (uint32_t)run, c3f};
            orc_philox4x32_10(ctr, key, sb);
        }
        const double lam = (p->D == 0.0 || c.n_live + filled == N_int)
                               ? 1e-20 : (p->D / p->D0) * (double)(N_int - (c.n_live + filled));
        const double dt_fill = (1.0 / lam) * -orc_log(((double)sb[2] + 0.5) * 0x1p-32);
        double dt_rec = dt_fill;
        if (c.n_live > 0) dt_rec = K > 0.0 ? -orc_log(((double)sb[0] + 0.5) * 0x1p-32) / K : INFINITY;
        const double dt = fmin(fmin(dt_fill, dt_rec), dt_cap);
        t += dt;
        uint8_t ev = 0;
        if (dt == dt_fill) {
            filled += 1;
            if (!(t >= duration && duration != 0.0)) { status = INT32_MIN + 1; break; }
        } else if (dt == dt_rec) {
            const double target = (double)sb[1] * 0x1p-32 * K;
            int pw = -1;
            double carry = 0.0;
            for (int r = 0; r < rows && pw < 0; ++r) {
                const double c2 = carry + rtot[r];
                if (c2 > target) {
                    ev_row(&R, r, v);
                    for (int l = 0; l < 64; ++l)
                        if (carry + v[l] > target) { pw = r * 64 + l; break; }
                    break;
                }
                carry = c2;
            }
            if (pw < 0)
                for (int s2 = c.n_live - 1; s2 >= 0 && pw < 0; --s2)
                    if (ev_rate(&R, s2) > 0.0) pw = s2;
            const int ew_ = c.live[pw];
            const uint32_t hw = c.nn[ew_];
            c.h_alive[hw] = 0;
            memmove(c.live + pw, c.live + pw + 1, sizeof(int32_t) * (size_t)(c.n_live - pw - 1));
            c.n_live -= 1;
            for (int s2 = 0; s2 < c.n_live; ++s2) {
                const int e = c.live[s2];
                if (c.nn[e] == hw) refresh_nn(&c, e, alpha_f, alpha_nm, 0);
            }
            ev = 1;
        }
        lum[i] = ev;
        if (tax) tax[i] = t;
        if (erat) erat[i] = (double)(c.n_live + filled) / p->N_e;
        ++i;
        if (duration != 0.0 && t >= duration) break;
    }
    if (live_steps) *live_steps = ls;
    free(rtot);
    cluster_free(&c);
    return status != 0 ? status : i;
}

int orc_run_event(const mcl_run_params* runs, int32_t n_runs, const double* duration,
                  uint64_t seed, int32_t max_steps, int32_t n_threads, uint8_t* lum,
                  double* t_axis, double* e_ratio, int32_t* steps)
{
    if (!runs || n_runs <= 0 || !lum || !steps || !duration || max_steps <= 0) return MCL_EINVAL;
    int64_t total = 0;
    for (int r = 0; r < n_runs; ++r) total += runs[r].sims;
    int32_t* c_run = xcalloc((size_t)total, sizeof(int32_t));
    int32_t* c_sim = xcalloc((size_t)total, sizeof(int32_t));
    int64_t k = 0;
    for (int r = 0; r < n_runs; ++r)
        for (int s = 0; s < runs[r].sims; ++s) { c_run[k] = r; c_sim[k] = s; ++k; }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < total; ++i) {
        const size_t o = (size_t)i * (size_t)max_steps;
        steps[i] = run_cluster_event(&runs[c_run[i]], c_run[i], c_sim[i], key, duration[c_run[i]],
                                     max_steps, lum + o, t_axis ? t_axis + o : NULL,
                                     e_ratio ? e_ratio + o : NULL, NULL);
    }
    free(c_run); free(c_sim);
    return 0;
}

/* --------------------------------------------------------- lab experiments -- */
/* Replay of the lab kernel (mcluminescence_amd/csrc/mc_lab.hip): TL_lab / ISO_lab
 * loops (tl_trap_lab.py:65-179) with dose filling, electrons and holes in creation
 * order with explicit float32 coordinates, nearest holes by full scan. */
typedef struct {
    float *hx, *hy, *hz;      /* hole slots; x = 3e38 marks a removed hole */
    float *ex, *ey, *ez, *ew;
    uint32_t* enn;
    int n_e, n_hs, h_cap, e_cap;
} lab_t;

#define LAB_DEAD 3e38f

static uint64_t lab_nearest(const lab_t* L, int n, float x, float y, float z)
{
    uint64_t key = ~0ull;
    for (int s = 0; s < n; ++s) {
        float d2;
        if (L->hx[s] == LAB_DEAD) d2 = INFINITY;
        else {
            float dx = L->hx[s] - x, dy = L->hy[s] - y, dz = L->hz[s] - z;
            d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        }
        uint32_t b;
        memcpy(&b, &d2, 4);
        const uint64_t k = ((uint64_t)b << 32) | (uint32_t)s;
        if (k < key) key = k;
    }
    return key;
}

static void lab_from_key(uint64_t key, float alpha_f, uint32_t* slot, float* w)
{
    const uint32_t b = (uint32_t)(key >> 32);
    if (b >= 0x7f800000u) { *slot = MCL_NO_HOLE; *w = 0.0f; return; }
    float d2;
    memcpy(&d2, &b, 4);
    *slot = (uint32_t)(key & 0xFFFFu);
    *w = orc_expf(-(alpha_f * sqrtf(d2)));
}

#define LAB_EXACT_FILL 0x1u
#define LAB_SHARED_CHANNEL 0x2u
#define LAB_RETRAP 0x4u
#define LAB_FIRST_ELECTRON 0x8u
#define LAB_LEGACY_GEOMETRY 0x10u

/* One electron-hole pair appended (Box.add_electron, engine.py:133-152); returns
 * 0, or -2 when the cluster is full.  exact: every nearest hole exact after it. */
static int lab_add_pair(lab_t* L, int* fills, int exact, int run, uint32_t c3e, uint32_t c3h,
                        const uint32_t key[2], double side_nm, double side_b_nm, float alpha_f,
                        uint32_t* remap)
{
    if (L->n_e >= L->e_cap) return -2;
    if (L->n_hs >= L->h_cap) {
        int base = 0;
        for (int s = 0; s < L->n_hs; ++s) {
            if (L->hx[s] == LAB_DEAD) { remap[s] = MCL_NO_HOLE; continue; }
            remap[s] = (uint32_t)base;
            L->hx[base] = L->hx[s]; L->hy[base] = L->hy[s]; L->hz[base] = L->hz[s];
            ++base;
        }
        for (int e = 0; e < L->n_e; ++e) if (L->enn[e] != MCL_NO_HOLE) L->enn[e] = remap[L->enn[e]];
        L->n_hs = base;
        if (L->n_hs >= L->h_cap) return -2;
    }
    uint32_t oe[4], oh[4];
    { uint32_t ctr[4] = {(uint32_t)*fills, 0xFFFFFFFEu, (uint32_t)run, c3e}; orc_philox4x32_10(ctr, key, oe); }
    { uint32_t ctr[4] = {(uint32_t)*fills, 0xFFFFFFFEu, (uint32_t)run, c3h}; orc_philox4x32_10(ctr, key, oh); }
    const float ex = (float)(((double)oe[0] * 0x1p-32) * side_nm), ey = (float)(((double)oe[1] * 0x1p-32) * side_nm),
                ez = (float)(((double)oe[2] * 0x1p-32) * side_nm);
    const int hs = L->n_hs;
    L->hx[hs] = (float)(((double)oh[0] * 0x1p-32) * side_b_nm);
    L->hy[hs] = (float)(((double)oh[1] * 0x1p-32) * side_b_nm);
    L->hz[hs] = (float)(((double)oh[2] * 0x1p-32) * side_b_nm);
    uint32_t slot;
    float w;
    lab_from_key(lab_nearest(L, exact ? hs + 1 : hs, ex, ey, ez), alpha_f, &slot, &w);
    if (exact) {
        for (int e = 0; e < L->n_e; ++e) {
            float dx = L->hx[hs] - L->ex[e], dy = L->hy[hs] - L->ey[e], dz = L->hz[hs] - L->ez[e];
            float d2n = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            float d2o = INFINITY;
            const uint32_t s0 = L->enn[e];
            if (s0 != MCL_NO_HOLE) {
                float ox = L->hx[s0] - L->ex[e], oy = L->hy[s0] - L->ey[e], oz = L->hz[s0] - L->ez[e];
                d2o = fmaf(oz, oz, fmaf(oy, oy, ox * ox));
            }
            if (d2n < d2o) { L->enn[e] = (uint32_t)hs; L->ew[e] = orc_expf(-(alpha_f * sqrtf(d2n))); }
        }
    }
    L->ex[L->n_e] = ex; L->ey[L->n_e] = ey; L->ez[L->n_e] = ez; L->ew[L->n_e] = w; L->enn[L->n_e] = slot;
    L->n_e += 1; L->n_hs += 1; *fills += 1;
    return 0;
}

static int32_t run_cluster_lab(const mcl_run_params* p, int run, int sim, const uint32_t key[2],
                               const double* obs, int n_obs, uint32_t options, int64_t max_steps,
                               int h_cap, int e_cap, double* out)
{
    mcl_geometry g;
    if (options & LAB_LEGACY_GEOMETRY) orc_run_geometry_legacy(p, &g);
    else orc_run_geometry(p, &g);
    const int exact_fill = (options & LAB_EXACT_FILL) != 0;
    const int shared_ch = (options & LAB_SHARED_CHANNEL) != 0;
    const int first_e = (options & LAB_FIRST_ELECTRON) != 0;
    const int retrap = (options & LAB_RETRAP) != 0;
    const uint64_t retrap_thr = p->Retrap >= 1.0 ? (1ull << 32)
                              : (p->Retrap > 0.0 ? (uint64_t)(p->Retrap * 0x1p32) : 0u);
    const double side_nm = g.side_m * 1e9, side_b_nm = (g.side_m * p->boundary_factor) * 1e9;
    lab_t L = {0};
    L.h_cap = h_cap; L.e_cap = e_cap;
    L.hx = xcalloc(h_cap, 4); L.hy = xcalloc(h_cap, 4); L.hz = xcalloc(h_cap, 4);
    L.ex = xcalloc(e_cap, 4); L.ey = xcalloc(e_cap, 4); L.ez = xcalloc(e_cap, 4);
    L.ew = xcalloc(e_cap, 4); L.enn = xcalloc(e_cap, 4);
    uint32_t* remap = xcalloc(h_cap, 4);
    const uint32_t c3e = ((uint32_t)sim << 4) | MCL_TAG_GEOM_E, c3h = ((uint32_t)sim << 4) | MCL_TAG_GEOM_H;
    const float alpha_f = (float)(p->alpha * 1e-9);
    L.n_e = g.n_e0;
    L.n_hs = g.n_h;
    for (int h = 0; h < L.n_hs; ++h) {
        uint32_t ctr[4] = {0u, (uint32_t)h, (uint32_t)run, c3h}, o[4];
        orc_philox4x32_10(ctr, key, o);
        L.hx[h] = (float)(((double)o[0] * 0x1p-32) * side_b_nm);
        L.hy[h] = (float)(((double)o[1] * 0x1p-32) * side_b_nm);
        L.hz[h] = (float)(((double)o[2] * 0x1p-32) * side_b_nm);
    }
    for (int e = 0; e < L.n_e; ++e) {
        uint32_t ctr[4] = {0u, (uint32_t)e, (uint32_t)run, c3e}, o[4];
        orc_philox4x32_10(ctr, key, o);
        L.ex[e] = (float)(((double)o[0] * 0x1p-32) * side_nm);
        L.ey[e] = (float)(((double)o[1] * 0x1p-32) * side_nm);
        L.ez[e] = (float)(((double)o[2] * 0x1p-32) * side_nm);
    }
    for (int e = 0; e < L.n_e; ++e)
        lab_from_key(lab_nearest(&L, L.n_hs, L.ex[e], L.ey[e], L.ez[e]), alpha_f, &L.enn[e], &L.ew[e]);

    const int two_ch = (p->E_loc_1 != p->E_loc_2) && (p->Retrap > 0.0) && (p->Retrap < 1.0);
    const int always2 = p->Retrap >= 1.0;
    const uint64_t ret_thr = two_ch ? (uint64_t)(p->Retrap * 0x1p32) : 0u;
    const int b1_is_a = p->b == p->s && p->E_loc_1 == p->E_cb;
    const int b2_is_b1 = p->E_loc_2 == p->E_loc_1;
    const uint32_t c3f = ((uint32_t)sim << 4) | MCL_TAG_FIRE, c3c = ((uint32_t)sim << 4) | MCL_TAG_CHAN;
    const int N_int = (int)p->N_e;
    const double T0_K = p->T_start + 273.15;
    double* rtot = xcalloc((size_t)(e_cap / 64 + 1), sizeof(double));
    double v[64];
    double t = 0.0;
    int64_t i = 0;
    int k_obs = 0, fills = 0;
    int32_t status = 0;
    while (k_obs < n_obs) {
        if (i >= max_steps) { status = -1; break; }
        const double T = T0_K + p->T_rate * t;
        const double kT = p->k_b * T;
        const double A = p->s * orc_exp(-(p->E_cb / kT));
        const double B1 = b1_is_a ? A : p->b * orc_exp(-(p->E_loc_1 / kT));
        const double B2 = b2_is_b1 ? B1 : p->b * orc_exp(-(p->E_loc_2 / kT));
        const int rows = (L.n_e + 63) / 64;
        double K = 0.0;
        /* the shared draw: one word per step for every electron (functions.py:125) */
        const int ch_step = (two_ch && shared_ch)
            ? (uint64_t)philox_word((uint32_t)(i >> 2), 0xFFFFFFFDu, (uint32_t)run, c3c, key, (int)(i & 3)) < ret_thr
            : always2;
#define LAB_RATE(pp) ((pp) >= L.n_e || L.enn[(pp)] == MCL_NO_HOLE ? 0.0 : \
        fma(((two_ch && !shared_ch ? (uint64_t)philox_word((uint32_t)(i >> 2), (uint32_t)(pp), (uint32_t)run, c3c, key, (int)(i & 3)) < ret_thr : ch_step) ? B2 : B1), \
            (double)L.ew[(pp)], A))
        for (int r = 0; r < rows; ++r) {
            for (int l = 0; l < 64; ++l) v[l] = LAB_RATE(r * 64 + l);
            scan64(v);
            rtot[r] = v[63];
            K = K + v[63];
        }
        uint32_t sb[4];
        { uint32_t ctr[4] = {(uint32_t)i, 0xFFFFFFFFu, (uint32_t)run, c3f}; orc_philox4x32_10(ctr, key, sb); }
        const double lam = (p->D == 0.0 || L.n_e == N_int) ? 1e-20 : (p->D / p->D0) * (double)(N_int - L.n_e);
        const double dt_fill = (1.0 / lam) * -orc_log(((double)sb[2] + 0.5) * 0x1p-32);
        double dt_rec = dt_fill;
        if (L.n_e > 0) dt_rec = K > 0.0 ? -orc_log(((double)sb[0] + 0.5) * 0x1p-32) / K : INFINITY;
        const double dt = fmin(dt_rec, dt_fill);
        t += dt;
        if (dt == dt_fill) {
            if (lab_add_pair(&L, &fills, exact_fill, run, c3e, c3h, key, side_nm, side_b_nm, alpha_f, remap)) {
                status = -2;
                break;
            }
        } else if (dt == dt_rec) {
            const double target = (double)sb[1] * 0x1p-32 * K;
            int pw = -1;
            double carry = 0.0;
            if (first_e)      /* recomber as the legacy loops call it: the oldest electron */
                for (int s2 = 0; s2 < L.n_e && pw < 0; ++s2) if (L.enn[s2] != MCL_NO_HOLE) pw = s2;
            for (int r = 0; r < rows && pw < 0; ++r) {
                const double c2 = carry + rtot[r];
                if (c2 > target) {
                    for (int l = 0; l < 64; ++l) v[l] = LAB_RATE(r * 64 + l);
                    scan64(v);
                    for (int l = 0; l < 64; ++l) if (carry + v[l] > target) { pw = r * 64 + l; break; }
                    break;
                }
                carry = c2;
            }
            if (pw < 0)
                for (int s2 = L.n_e - 1; s2 >= 0 && pw < 0; --s2) if (LAB_RATE(s2) > 0.0) pw = s2;
#undef LAB_RATE
            const uint32_t hw = L.enn[pw];
            L.hx[hw] = LAB_DEAD;
            for (int q = pw + 1; q < L.n_e; ++q) {
                L.ex[q - 1] = L.ex[q]; L.ey[q - 1] = L.ey[q]; L.ez[q - 1] = L.ez[q];
                L.ew[q - 1] = L.ew[q]; L.enn[q - 1] = L.enn[q];
            }
            L.n_e -= 1;
            for (int e = 0; e < L.n_e; ++e)
                if (L.enn[e] == hw)
                    lab_from_key(lab_nearest(&L, L.n_hs, L.ex[e], L.ey[e], L.ez[e]), alpha_f, &L.enn[e], &L.ew[e]);
            /* legacy retrapping (functions.py:328-331): a new pair, exact nearest holes */
            if (retrap && (uint64_t)sb[3] < retrap_thr &&
                lab_add_pair(&L, &fills, 1, run, c3e, c3h, key, side_nm, side_b_nm, alpha_f, remap)) {
                status = -2;
                break;
            }
        }
        ++i;
        while (k_obs < n_obs && t >= obs[k_obs]) { out[k_obs] = (double)L.n_e / p->N_e; ++k_obs; }
    }
    free(rtot); free(remap);
    free(L.hx); free(L.hy); free(L.hz); free(L.ex); free(L.ey); free(L.ez); free(L.ew); free(L.enn);
    return status;
}

int orc_run_lab(const mcl_run_params* runs, int32_t n_runs, const double* obs_times,
                const int32_t* obs_count, uint32_t options, uint64_t seed, int64_t max_steps,
                int32_t n_threads, double* ratio, int32_t* status)
{
    if (!runs || n_runs <= 0 || !obs_times || !obs_count || !ratio) return MCL_EINVAL;
    int64_t total = 0;
    int max_obs = 0, max_e = 0, max_h = 0;
    int32_t* offs = xcalloc((size_t)n_runs, sizeof(int32_t));
    int64_t o = 0;
    for (int r = 0; r < n_runs; ++r) {
        offs[r] = (int32_t)o;
        o += obs_count[r];
        if (obs_count[r] > max_obs) max_obs = obs_count[r];
        total += runs[r].sims;
        if (runs[r].sims > 0) {
            mcl_geometry g;
            if (options & LAB_LEGACY_GEOMETRY) orc_run_geometry_legacy(&runs[r], &g);
            else orc_run_geometry(&runs[r], &g);
            int ne = g.n_e0 > (int)runs[r].N_e ? g.n_e0 : (int)runs[r].N_e;
            if (ne > max_e) max_e = ne;
            if (g.n_h > max_h) max_h = g.n_h;
        }
    }
    const int e_cap = ((max_e + 64 + 63) / 64) * 64;
    const int h_cap = ((max_h + 2 * max_e + 64 + 3) / 4) * 4;
    int32_t* c_run = xcalloc((size_t)total, sizeof(int32_t));
    int32_t* c_sim = xcalloc((size_t)total, sizeof(int32_t));
    int64_t k = 0;
    for (int r = 0; r < n_runs; ++r)
        for (int s = 0; s < runs[r].sims; ++s) { c_run[k] = r; c_sim[k] = s; ++k; }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t c = 0; c < total; ++c) {
        double* out = ratio + (size_t)c * (size_t)max_obs;
        memset(out, 0, sizeof(double) * (size_t)max_obs);
        const int r = c_run[c];
        const int32_t st = run_cluster_lab(&runs[r], r, c_sim[c], key, obs_times + offs[r],
                                           obs_count[r], options, max_steps, h_cap, e_cap, out);
        if (status) status[c] = st;
    }
    free(c_run); free(c_sim); free(offs);
    return 0;
}
