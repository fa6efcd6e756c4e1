// mlp_abi.hip — host side of the surrogate network (include/mcl_surrogate.h):
// parameter storage, batch buffers, the forward / backward launch sequence and
// the optimizer step, all on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <new>
#include <vector>

#include "../../include/mcl_surrogate.h"
#include "mcl_ctx.h"
#include "mlp_internal.h"

namespace {

using mcl::Ctx;
using mcl::fail;
namespace M = mcl::mlp;
typedef __bf16 bf16;

constexpr int kMaxSplits = 8;   // split-K factor cap of the weight-gradient products

struct Layer {
    int in = 0, out = 0, kp = 0;          // kp: inputs padded for the K tiling
    int64_t flat_w = 0, flat_b = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr;     // [out][kp]
    float *bw = nullptr, *bm = nullptr, *bv = nullptr;  // [out]
    bf16 *wb = nullptr, *wt = nullptr;                  // [out][kp], [kp][out]
    float* part = nullptr;                              // [kMaxSplits][out][kp]
};

struct Mlp {
    Ctx* ctx = nullptr;
    mcl_mlp_desc d{};
    std::vector<Layer> L;                 // n_hidden_layers + 1 linear layers
    int64_t n_params = 0;
    std::vector<void*> allocs;
    bf16 *X0 = nullptr, *X0t = nullptr;   // [B][kp0], [kp0][B]
    std::vector<bf16*> Z, A, At, dZ, dZt; // hidden layer l: [B][H] ([H][B] for *t)
    bf16 *dY = nullptr, *dYt = nullptr;   // [B][n_out], [n_out][B]
    float* loss_part = nullptr;
    float* loss_tmp = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    // the fused inference kernel's network: weights in fragment order (repacked
    // after any change of the weights), biases, LDS bytes; usable iff fused_ok
    M::DeNet fnet{};
    bf16* wf = nullptr;
    bool wf_stale = true;
    bool fused_ok = false;
    int fused_lds = 0;

    ~Mlp()
    {
        for (void* p : allocs) (void)hipFree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

template <typename T>
int dalloc(Mlp* m, T** p, int64_t count)
{
    void* q = nullptr;
    HIP_TRY(m->ctx, hipMalloc(&q, (size_t)count * sizeof(T)));
    m->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return MCL_OK;
}

#define TRY(expr)                        \
    do {                                 \
        int st_ = (expr);                \
        if (st_ != MCL_OK) return st_;   \
    } while (0)

// Largest tile that still gives every CU a workgroup (M, N multiples of 64).
int pick_tile(const Mlp* m, int Mr, int Nc)
{
    const int tiles[] = {M::kTile128x128, M::kTile128x64, M::kTile64x64};
    int best = -1;
    for (int t : tiles) {
        if (Mr % M::tile_m(t) || Nc % M::tile_n(t)) continue;
        if (best < 0) best = t;
        if ((int64_t)(Mr / M::tile_m(t)) * (Nc / M::tile_n(t)) >= m->ctx->n_cu) return t;
        best = t;
    }
    return best;
}

// Split-K factor of a weight-gradient product: enough workgroups for two per CU,
// every split a multiple of 64 rows of the batch.
int pick_splits(const Mlp* m, int tiles, int batch)
{
    int s = 1;
    while (s < kMaxSplits && (int64_t)tiles * s < 2 * m->ctx->n_cu && batch % (128 * s) == 0) s *= 2;
    return s;
}

int gemm(Mlp* m, const M::GemmArgs& g, int epi, int splits = 1)
{
    const int tile = pick_tile(m, g.M, g.N);
    if (tile < 0) return fail(m->ctx, MCL_EINVAL, "no tile shape divides a %d x %d product", g.M, g.N);
    if (M::gemm_resident_ok(g, tile, epi, splits))
        HIP_TRY(m->ctx, M::gemm_resident(g, epi, m->ctx->stream));
    else
        HIP_TRY(m->ctx, M::gemm(g, tile, epi, splits, m->ctx->stream));
    return MCL_OK;
}

int check_batch(Mlp* m, int32_t batch)
{
    if (batch <= 0 || batch % 64 || batch > m->d.max_batch)
        return fail(m->ctx, MCL_EINVAL, "batch %d must be a positive multiple of 64 and at most %d",
                    batch, m->d.max_batch);
    return MCL_OK;
}

M::OptArgs opt_args(Mlp* m, int mode, int batch, int* groups)
{
    M::OptArgs a{};
    a.n_layers = (int)m->L.size();
    a.mode = mode;
    a.batch = batch;
    int g = 0;
    for (size_t i = 0; i < m->L.size(); ++i) {
        const Layer& s = m->L[i];
        M::OptLayer& o = a.L[i];
        o.w = s.w; o.m = s.m; o.v = s.v; o.wb = s.wb; o.wt = s.wt;
        o.part = s.part; o.splits = 0;
        o.bw = s.bw; o.bm = s.bm; o.bv = s.bv;
        o.dzt = i + 1 < m->L.size() ? m->dZt[i] : m->dYt;
        o.out = s.out; o.kp = s.kp; o.k_in = s.in;
        o.flat_w = s.flat_w; o.flat_b = s.flat_b;
        o.tile0 = g;
        g += M::opt_tiles(s.out, s.kp);
    }
    for (size_t i = 0; i < m->L.size(); ++i) {
        a.L[i].bias0 = g;
        g += M::opt_bias_groups(m->L[i].out);
    }
    *groups = g;
    return a;
}

void begin(Mlp* m)
{
    m->timed = hipEventRecord(m->ev0, m->ctx->stream) == hipSuccess;
}

int end(Mlp* m)
{
    if (m->timed) m->timed = hipEventRecord(m->ev1, m->ctx->stream) == hipSuccess;
    return MCL_OK;
}

// The layers of the forward pass on the packed input m->X0 (batch B, a multiple
// of 64), on the context's stream.
int forward_packed(Mlp* m, int32_t B, float* y)
{
    const bf16* in = m->X0;
    const int nl = (int)m->L.size();
    for (int l = 0; l < nl; ++l) {
        const Layer& s = m->L[l];
        M::GemmArgs g{};
        g.A = in; g.lda = s.kp;
        g.B = s.wb; g.ldb = s.kp;
        g.M = B; g.N = s.out; g.K = s.kp;
        g.bias = s.bw;
        if (l + 1 < nl) {
            g.out_a = m->A[l % 2];
            TRY(gemm(m, g, M::kHiddenInfer));
            in = m->A[l % 2];
        } else {
            g.out_f = y;
            TRY(gemm(m, g, M::kOutInfer));
        }
    }
    return MCL_OK;
}

// The forward pass of batch B: one kernel for the whole network where its shape
// allows (the context's "mlp_fused" knob, default on), else pack x into m->X0 and
// run the layers; the two give identical outputs.
int forward(Mlp* m, const float* x, int32_t B, float* y)
{
    hipStream_t st = m->ctx->stream;
    if (m->fused_ok && m->ctx->mlp_fused != 0.0) {
        if (m->wf_stale) {
            int64_t off = 0;
            for (size_t l = 0; l < m->L.size(); ++l) {
                const Layer& s = m->L[l];
                HIP_TRY(m->ctx, M::de_pack(s.wb, s.out, s.kp, m->wf + off, st));
                off += (int64_t)s.out * s.kp;
            }
            m->wf_stale = false;
        }
        HIP_TRY(m->ctx, M::infer_fused(x, B, m->d.n_in, m->fnet, y, m->d.n_out, m->fused_lds, st));
        return MCL_OK;
    }
    HIP_TRY(m->ctx, M::pack_input(x, nullptr, m->d.n_in, B, m->L[0].kp, m->X0, nullptr, st));
    return forward_packed(m, B, y);
}

// Network-scored differential evolution: the population (two buffers: the one a
// generation reads and the one its selection writes) and the network's input /
// output rows of one generation, in device memory.
struct De {
    Mlp* m = nullptr;
    M::DeArgs a{};
    float* xb[2] = {nullptr, nullptr};                  // [pop][dim]
    unsigned long long* kb[2] = {nullptr, nullptr};     // argmin keys [pop]
    int cur = 0;                          // the buffer holding the population
    uint32_t gen = 0;                     // generations run (0: the initial population)
    bool fused = false;                   // one kernel per generation (de_generation_kernel)
    M::DeNet net{};
    int fused_lds = 0;
    std::vector<void*> allocs;

    ~De()
    {
        for (void* p : allocs) (void)hipFree(p);
    }
};

// Points the arguments at the current population and the one selection writes.
void de_buffers(De* d)
{
    d->a.x = d->xb[d->cur];
    d->a.x_next = d->xb[d->cur ^ 1];
    d->a.keys = d->kb[d->cur];
    d->a.keys_next = d->kb[d->cur ^ 1];
}

// One generation on the context's stream: trials, their network curves, selection.
int de_generation(De* d)
{
    Mlp* m = d->m;
    hipStream_t st = m->ctx->stream;
    ++d->gen;
    de_buffers(d);
    if (d->fused) {
        HIP_TRY(m->ctx, M::de_generation_fused(d->a, d->net, d->gen, d->fused_lds, st));
    } else {
        HIP_TRY(m->ctx, M::de_propose(d->a, d->gen, st));   // also packs the trials into X0
        TRY(forward_packed(m, d->a.pop_pad, const_cast<float*>(d->a.Y)));
        HIP_TRY(m->ctx, M::de_score(d->a, st));
    }
    d->cur ^= 1;
    return MCL_OK;
}

// Whether the fused generation kernel takes the network (every layer at most 512
// outputs, K a multiple of 64 up to 512, its LDS within a workgroup's); fills
// net with the layers' shapes and biases (weights: set by de_fused_weights).
bool de_fused_net(const Mlp* m, M::DeNet* net, int* lds)
{
    if (m->L.size() > (size_t)M::kMaxLayers) return false;
    int widest = 0;
    for (size_t l = 0; l < m->L.size(); ++l) {
        const Layer& s = m->L[l];
        if (s.out % 32 || s.out > 512 || s.kp % 64 || s.kp > 512) return false;
        net->L[l] = M::DeLayer{nullptr, s.bw, s.out, s.kp};
        widest = std::max(widest, s.kp);
        if (l + 1 < m->L.size()) widest = std::max(widest, s.out);
    }
    net->n_layers = (int)m->L.size();
    net->lda = widest + 8;
    *lds = M::de_fused_lds(*net, m->d.n_out);
    return *lds <= m->ctx->max_lds;
}

// The fused kernel's copy of the weights in fragment order (one buffer for all
// layers), packed on the context's stream from the network's current weights.
int de_fused_weights(De* d, const std::function<hipError_t(void**, size_t)>& alloc)
{
    Mlp* m = d->m;
    size_t total = 0;
    for (const Layer& s : m->L) total += (size_t)s.out * s.kp;
    bf16* wf = nullptr;
    const hipError_t e = alloc((void**)&wf, total * sizeof(bf16));
    if (e != hipSuccess)
        return fail(m->ctx, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE,
                    "fused DE weights: %s", hipGetErrorString(e));
    size_t off = 0;
    for (size_t l = 0; l < m->L.size(); ++l) {
        const Layer& s = m->L[l];
        HIP_TRY(m->ctx, M::de_pack(s.wb, s.out, s.kp, wf + off, m->ctx->stream));
        d->net.L[l].w = wf + off;
        off += (size_t)s.out * s.kp;
    }
    return MCL_OK;
}

// Forward + backward of the batch: gradient partial sums in every layer's `part`
// (splits recorded in *splits), bias gradients as row sums of dZ^T, loss in *loss.
int forward_backward(Mlp* m, const float* x, const float* t, const int32_t* idx, int32_t B,
                     float* loss, std::vector<int>& splits)
{
    const int nl = (int)m->L.size(), nh = nl - 1;
    hipStream_t st = m->ctx->stream;
    HIP_TRY(m->ctx, M::pack_input(x, idx, m->d.n_in, B, m->L[0].kp, m->X0, m->X0t, st));
    const bf16* in = m->X0;
    for (int l = 0; l < nh; ++l) {
        M::GemmArgs g{};
        g.A = in; g.lda = m->L[l].kp;
        g.B = m->L[l].wb; g.ldb = m->L[l].kp;
        g.M = B; g.N = m->L[l].out; g.K = m->L[l].kp;
        g.bias = m->L[l].bw;
        g.out_z = m->Z[l]; g.out_a = m->A[l]; g.out_t = m->At[l]; g.ldt = B;
        TRY(gemm(m, g, M::kHiddenTrain));
        in = m->A[l];
    }
    const Layer& lo = m->L[nh];
    M::GemmArgs g{};
    g.A = in; g.lda = lo.kp;
    g.B = lo.wb; g.ldb = lo.kp;
    g.M = B; g.N = lo.out; g.K = lo.kp;
    g.bias = lo.bw;
    g.out_a = m->dY; g.out_t = m->dYt; g.ldt = B;
    g.target = t; g.t_idx = idx; g.ld_target = lo.out;
    g.grad_scale = (float)(2.0 / ((double)B * lo.out));
    g.loss_part = m->loss_part;
    TRY(gemm(m, g, M::kOutTrain));
    const int tile = pick_tile(m, B, lo.out);
    const int n_part = (B / M::tile_m(tile)) * (lo.out / M::tile_n(tile));
    HIP_TRY(m->ctx, M::finish_loss(m->loss_part, n_part, (float)(1.0 / ((double)B * lo.out)),
                                   loss ? loss : m->loss_tmp, st));

    splits.assign(nl, 1);
    for (int l = nl - 1; l >= 0; --l) {
        const Layer& s = m->L[l];
        const bf16* dz = l == nh ? m->dY : m->dZ[l];
        const bf16* dzt = l == nh ? m->dYt : m->dZt[l];
        const bf16* xt = l == 0 ? m->X0t : m->At[l - 1];
        // weight gradient dW = dZ^T X, split over the batch
        M::GemmArgs w{};
        w.A = dzt; w.lda = B;
        w.B = xt; w.ldb = B;
        w.M = s.out; w.N = s.kp;
        const int tile_w = pick_tile(m, s.out, s.kp);
        if (tile_w < 0) return fail(m->ctx, MCL_EINVAL, "no tile shape for layer %d", l);
        const int tiles = (s.out / M::tile_m(tile_w)) * (s.kp / M::tile_n(tile_w));
        splits[l] = pick_splits(m, tiles, B);
        w.K = B / splits[l];
        w.out_f = s.part;
        TRY(gemm(m, w, M::kDwPartial, splits[l]));
        if (l == 0) break;
        // input gradient through the SiLU of layer l-1: dZ_{l-1} = (dZ_l W_l) * silu'(Z_{l-1})
        M::GemmArgs x2{};
        x2.A = dz; x2.lda = s.out;
        x2.B = s.wt; x2.ldb = s.out;
        x2.M = B; x2.N = s.kp; x2.K = s.out;
        x2.z_in = m->Z[l - 1];
        x2.out_a = l - 1 >= 1 ? m->dZ[l - 1] : nullptr;
        x2.out_t = m->dZt[l - 1]; x2.ldt = B;
        TRY(gemm(m, x2, M::kDxSilu));
    }
    return MCL_OK;
}

int run_optimizer(Mlp* m, int mode, int B, const std::vector<int>& splits, float* flat,
                  const mcl_adamw* hp)
{
    int groups = 0;
    M::OptArgs a = opt_args(m, mode, B, &groups);
    for (size_t i = 0; i < m->L.size(); ++i) a.L[i].splits = i < splits.size() ? splits[i] : 0;
    a.flat = flat;
    if (hp) {
        const double t = (double)hp->step;
        a.lr = (float)hp->lr;
        a.beta1 = (float)hp->beta1;
        a.beta2 = (float)hp->beta2;
        a.eps = (float)hp->eps;
        a.wd = (float)hp->weight_decay;
        a.step_size = (float)(hp->lr / (1.0 - std::pow(hp->beta1, t)));
        a.sqrt_bc2 = (float)std::sqrt(1.0 - std::pow(hp->beta2, t));
    }
    HIP_TRY(m->ctx, M::optimizer(a, groups, m->ctx->stream));
    return MCL_OK;
}

}  // namespace

extern "C" {

int mcl_mlp_create(void* ctxp, const mcl_mlp_desc* desc, void** out)
{
    Ctx* c = static_cast<Ctx*>(ctxp);
    if (!c || !desc || !out) return fail(c, MCL_EINVAL, "null argument");
    *out = nullptr;
    const mcl_mlp_desc& d = *desc;
    if (d.n_in < 1 || d.n_in > 64) return fail(c, MCL_EINVAL, "n_in %d not in [1, 64]", d.n_in);
    if (d.n_hidden < 128 || d.n_hidden % 128)
        return fail(c, MCL_EINVAL, "n_hidden %d must be a positive multiple of 128", d.n_hidden);
    if (d.n_out < 128 || d.n_out % 128)
        return fail(c, MCL_EINVAL, "n_out %d must be a positive multiple of 128", d.n_out);
    if (d.n_hidden_layers < 1 || d.n_hidden_layers + 1 > M::kMaxLayers)
        return fail(c, MCL_EINVAL, "n_hidden_layers %d not in [1, %d]", d.n_hidden_layers,
                    M::kMaxLayers - 1);
    if (d.max_batch < 64 || d.max_batch % 64)
        return fail(c, MCL_EINVAL, "max_batch %d must be a positive multiple of 64", d.max_batch);
    HIP_TRY(c, hipSetDevice(c->device));
    Mlp* m = new (std::nothrow) Mlp;
    if (!m) return fail(c, MCL_ENOMEM, "host allocation failed");
    m->ctx = c;
    m->d = d;
    const int nl = d.n_hidden_layers + 1;
    int64_t off = 0;
    for (int l = 0; l < nl; ++l) {
        Layer s;
        s.in = l == 0 ? d.n_in : d.n_hidden;
        s.out = l == nl - 1 ? d.n_out : d.n_hidden;
        s.kp = l == 0 ? 64 : s.in;
        s.flat_w = off;
        off += (int64_t)s.out * s.in;
        s.flat_b = off;
        off += s.out;
        m->L.push_back(s);
    }
    m->n_params = off;
    const int64_t B = d.max_batch;
    int st = MCL_OK;
    auto run = [&]() -> int {
        for (Layer& s : m->L) {
            const int64_t nw = (int64_t)s.out * s.kp;
            TRY(dalloc(m, &s.w, nw)); TRY(dalloc(m, &s.m, nw)); TRY(dalloc(m, &s.v, nw));
            TRY(dalloc(m, &s.bw, s.out)); TRY(dalloc(m, &s.bm, s.out)); TRY(dalloc(m, &s.bv, s.out));
            TRY(dalloc(m, &s.wb, nw)); TRY(dalloc(m, &s.wt, nw));
            TRY(dalloc(m, &s.part, kMaxSplits * nw));
            HIP_TRY(c, hipMemset(s.w, 0, nw * sizeof(float)));
            HIP_TRY(c, hipMemset(s.m, 0, nw * sizeof(float)));
            HIP_TRY(c, hipMemset(s.v, 0, nw * sizeof(float)));
            HIP_TRY(c, hipMemset(s.bw, 0, s.out * sizeof(float)));
            HIP_TRY(c, hipMemset(s.bm, 0, s.out * sizeof(float)));
            HIP_TRY(c, hipMemset(s.bv, 0, s.out * sizeof(float)));
            HIP_TRY(c, hipMemset(s.wb, 0, nw * sizeof(bf16)));
            HIP_TRY(c, hipMemset(s.wt, 0, nw * sizeof(bf16)));
        }
        const int H = d.n_hidden;
        TRY(dalloc(m, &m->X0, B * 64));
        TRY(dalloc(m, &m->X0t, B * 64));
        for (int l = 0; l < d.n_hidden_layers; ++l) {
            bf16 *z, *a, *at, *dz, *dzt;
            TRY(dalloc(m, &z, B * H)); TRY(dalloc(m, &a, B * H)); TRY(dalloc(m, &at, B * H));
            TRY(dalloc(m, &dz, B * H)); TRY(dalloc(m, &dzt, B * H));
            m->Z.push_back(z); m->A.push_back(a); m->At.push_back(at);
            m->dZ.push_back(dz); m->dZt.push_back(dzt);
        }
        if (m->A.size() < 2) {   // inference ping-pong needs two activation buffers
            bf16* a2;
            TRY(dalloc(m, &a2, B * H));
            m->A.push_back(a2);
        }
        TRY(dalloc(m, &m->dY, B * d.n_out));
        TRY(dalloc(m, &m->dYt, B * d.n_out));
        TRY(dalloc(m, &m->loss_part, (B / 64) * (d.n_out / 64)));
        TRY(dalloc(m, &m->loss_tmp, 1));
        HIP_TRY(c, hipEventCreate(&m->ev0));
        HIP_TRY(c, hipEventCreate(&m->ev1));
        // the fused inference kernel's view of the network
        int64_t nwf = 0, widest = 0;
        for (size_t l = 0; l < m->L.size(); ++l) {
            const Layer& s = m->L[l];
            m->fnet.L[l] = M::DeLayer{nullptr, s.bw, s.out, s.kp};
            nwf += (int64_t)s.out * s.kp;
            widest = std::max<int64_t>(widest, s.kp);
            if (l + 1 < m->L.size()) widest = std::max<int64_t>(widest, s.out);
        }
        m->fnet.n_layers = (int)m->L.size();
        m->fnet.lda = (int)widest + 8;
        m->fused_lds = M::infer_fused_lds(m->fnet);
        if (M::infer_fused_ok(m->fnet) && m->fused_lds <= c->max_lds &&
            M::infer_fused_prepare(m->fused_lds) == hipSuccess) {
            TRY(dalloc(m, &m->wf, nwf));
            int64_t off = 0;
            for (size_t l = 0; l < m->L.size(); ++l) {
                m->fnet.L[l].w = m->wf + off;
                off += (int64_t)m->L[l].out * m->L[l].kp;
            }
            m->fused_ok = true;
        }
        (void)hipGetLastError();
        return MCL_OK;
    };
    st = run();
    if (st != MCL_OK) {
        delete m;
        return st;
    }
    *out = m;
    return MCL_OK;
}

void mcl_mlp_destroy(void* mlp)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m) return;
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

int64_t mcl_mlp_param_count(const void* mlp)
{
    const Mlp* m = static_cast<const Mlp*>(mlp);
    return m ? m->n_params : -1;
}

int mcl_mlp_get_params(void* mlp, float* params)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !params) return MCL_EINVAL;
    HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
    std::vector<float> buf;
    for (const Layer& s : m->L) {
        buf.resize((size_t)s.out * s.kp);
        HIP_TRY(m->ctx, hipMemcpy(buf.data(), s.w, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int n = 0; n < s.out; ++n)
            std::copy(buf.begin() + (size_t)n * s.kp, buf.begin() + (size_t)n * s.kp + s.in,
                      params + s.flat_w + (int64_t)n * s.in);
        HIP_TRY(m->ctx, hipMemcpy(params + s.flat_b, s.bw, s.out * sizeof(float), hipMemcpyDeviceToHost));
    }
    return MCL_OK;
}

int mcl_mlp_set_params(void* mlp, const float* params)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !params) return MCL_EINVAL;
    Ctx* c = m->ctx;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<float> buf;
    for (const Layer& s : m->L) {
        buf.assign((size_t)s.out * s.kp, 0.0f);
        for (int n = 0; n < s.out; ++n)
            for (int k = 0; k < s.in; ++k) {
                const float v = params[s.flat_w + (int64_t)n * s.in + k];
                if (!std::isfinite(v)) return fail(c, MCL_EINVAL, "non-finite parameter");
                buf[(size_t)n * s.kp + k] = v;
            }
        const size_t nw = buf.size();
        HIP_TRY(c, hipMemcpy(s.w, buf.data(), nw * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(s.bw, params + s.flat_b, s.out * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemset(s.m, 0, nw * sizeof(float)));
        HIP_TRY(c, hipMemset(s.v, 0, nw * sizeof(float)));
        HIP_TRY(c, hipMemset(s.bm, 0, s.out * sizeof(float)));
        HIP_TRY(c, hipMemset(s.bv, 0, s.out * sizeof(float)));
    }
    TRY(run_optimizer(m, M::kOptRefresh, 64, {}, nullptr, nullptr));
    m->wf_stale = true;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MCL_OK;
}

int mcl_mlp_forward(void* mlp, const float* x, int32_t B, float* y)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !x || !y) return MCL_EINVAL;
    TRY(check_batch(m, B));
    begin(m);
    TRY(forward(m, x, B, y));
    return end(m);
}

int mcl_mlp_grad(void* mlp, const float* x, const float* t, const int32_t* idx, int32_t B,
                 float* grad, float* loss)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !x || !t) return MCL_EINVAL;
    TRY(check_batch(m, B));
    begin(m);
    std::vector<int> splits;
    TRY(forward_backward(m, x, t, idx, B, loss, splits));
    if (grad) TRY(run_optimizer(m, M::kOptExportGrad, B, splits, grad, nullptr));
    return end(m);
}

int mcl_mlp_train_step(void* mlp, const float* x, const float* t, const int32_t* idx, int32_t B,
                       const mcl_adamw* hp, float* loss)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !x || !t || !hp) return MCL_EINVAL;
    TRY(check_batch(m, B));
    if (hp->step < 1 || !(hp->lr >= 0) || !(hp->beta1 >= 0 && hp->beta1 < 1) ||
        !(hp->beta2 >= 0 && hp->beta2 < 1) || !(hp->eps > 0) || !(hp->weight_decay >= 0))
        return fail(m->ctx, MCL_EINVAL, "invalid AdamW hyper-parameters");
    begin(m);
    std::vector<int> splits;
    TRY(forward_backward(m, x, t, idx, B, loss, splits));
    TRY(run_optimizer(m, M::kOptAdamW, B, splits, nullptr, hp));
    m->wf_stale = true;
    return end(m);
}

int mcl_mlp_last_ms(void* mlp, double* ms)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !ms) return MCL_EINVAL;
    if (!m->timed) return fail(m->ctx, MCL_EINVAL, "no timed call");
    HIP_TRY(m->ctx, hipEventSynchronize(m->ev1));
    float f = 0.0f;
    HIP_TRY(m->ctx, hipEventElapsedTime(&f, m->ev0, m->ev1));
    *ms = f;
    return MCL_OK;
}

int mcl_de_create(void* mlp, const mcl_de_params* p, const float* target, void** de_out)
{
    Mlp* m = static_cast<Mlp*>(mlp);
    if (!m || !p || !target || !de_out) return MCL_EINVAL;
    *de_out = nullptr;
    const int n_in = m->d.n_in, n_out = m->d.n_out;
    const int pad = ((p->pop + 63) / 64) * 64;
    if (p->pop < 4 || pad > m->d.max_batch)
        return fail(m->ctx, MCL_EINVAL, "population %d must be at least 4 and, rounded up to 64, "
                    "at most the network's max_batch %d", p->pop, m->d.max_batch);
    if (p->dim < 1 || p->dim > 8 || p->dim > n_in || n_in - p->dim > 8)
        return fail(m->ctx, MCL_EINVAL, "dim %d: 1..8 searched inputs of %d, at most 8 fixed",
                    p->dim, n_in);
    if (!(p->F >= 0.0f && p->F <= 2.0f) || !(p->CR >= 0.0f && p->CR <= 1.0f))
        return fail(m->ctx, MCL_EINVAL, "F %g must be in [0, 2] and CR %g in [0, 1]", p->F, p->CR);
    if (!std::isfinite(p->log_span) || !std::isfinite(p->out_scale) || !(p->inv_peak2 >= 0.0f))
        return fail(m->ctx, MCL_EINVAL, "log_span, out_scale and inv_peak2 must be finite");
    if (p->flags & ~MCL_DE_LAYERED)
        return fail(m->ctx, MCL_EINVAL, "flags 0x%x: unknown bits", (unsigned)p->flags);
    De* d = new (std::nothrow) De();
    if (!d) return fail(m->ctx, MCL_ENOMEM, "host allocation failed");
    d->m = m;
    auto alloc = [&](void** q, size_t bytes) -> hipError_t {
        const hipError_t e = hipMalloc(q, bytes);
        if (e == hipSuccess) d->allocs.push_back(*q);
        return e;
    };
    M::DeArgs& a = d->a;
    float* Y = nullptr;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = alloc((void**)&d->xb[b], sizeof(float) * (size_t)p->pop * p->dim);
        if (e == hipSuccess) e = alloc((void**)&d->kb[b], sizeof(unsigned long long) * (size_t)p->pop);
    }
    if (e == hipSuccess) e = alloc((void**)&a.f, sizeof(float) * (size_t)p->pop);
    if (e == hipSuccess) e = alloc((void**)&a.X, sizeof(float) * (size_t)pad * n_in);
    if (e == hipSuccess) e = alloc((void**)&Y, sizeof(float) * (size_t)pad * n_out);
    if (e != hipSuccess) {
        delete d;
        return fail(m->ctx, e == hipErrorOutOfMemory ? MCL_ENOMEM : MCL_EDEVICE, "DE buffers: %s",
                    hipGetErrorString(e));
    }
    a.Y = Y;
    a.X0 = m->X0;
    a.kp0 = m->L[0].kp;
    a.target = target;
    a.pop = p->pop;
    a.pop_pad = pad;
    a.dim = p->dim;
    a.n_in = n_in;
    a.n_out = n_out;
    a.F = p->F;
    a.CR = p->CR;
    a.log_span = p->log_span;
    a.out_scale = p->out_scale;
    a.inv_peak2 = p->inv_peak2;
    a.key0 = (uint32_t)p->seed;
    a.key1 = (uint32_t)(p->seed >> 32);
    for (int k = 0; k < 8; ++k) a.fixed[k] = k < n_in - p->dim ? p->fixed[k] : 0.0f;
    d->fused = !(p->flags & MCL_DE_LAYERED) && de_fused_net(m, &d->net, &d->fused_lds);
    // generation 0: the initial population, uniform in [-1, 1], scored
    hipStream_t st = m->ctx->stream;
    int rc = MCL_OK;
    if (d->fused && M::de_fused_prepare(d->fused_lds) != hipSuccess)
        rc = fail(m->ctx, MCL_EDEVICE, "fused generation kernel: LDS attribute");
    if (rc == MCL_OK && d->fused) rc = de_fused_weights(d, alloc);
    de_buffers(d);
    if (rc == MCL_OK && M::de_init(a, st) != hipSuccess) rc = fail(m->ctx, MCL_EDEVICE, "de_init launch");
    if (rc == MCL_OK) rc = forward(m, a.X, pad, Y);
    if (rc == MCL_OK && M::de_score(a, st) != hipSuccess)
        rc = fail(m->ctx, MCL_EDEVICE, "de_score launch");
    d->cur ^= 1;
    if (rc != MCL_OK) {
        (void)hipStreamSynchronize(st);
        delete d;
        return rc;
    }
    *de_out = d;
    return MCL_OK;
}

void mcl_de_destroy(void* de)
{
    De* d = static_cast<De*>(de);
    if (!d) return;
    (void)hipStreamSynchronize(d->m->ctx->stream);
    delete d;
}

int mcl_de_run(void* de, int32_t gens)
{
    De* d = static_cast<De*>(de);
    if (!d) return MCL_EINVAL;
    if (gens < 0) return fail(d->m->ctx, MCL_EINVAL, "gens %d < 0", gens);
    for (int g = 0; g < gens; ++g) TRY(de_generation(d));
    return MCL_OK;
}

int mcl_de_best(void* de, float* x, float* f)
{
    De* d = static_cast<De*>(de);
    if (!d || !x || !f) return MCL_EINVAL;
    Ctx* c = d->m->ctx;
    // the members' argmin keys (objective bits << 32 | member: the objectives are
    // >= 0, so the smallest key is the first index of the minimum, unordered
    // objectives last) and the population, with one synchronisation
    const size_t pop = (size_t)d->a.pop, dim = (size_t)d->a.dim;
    std::vector<unsigned long long> kh(pop);
    std::vector<float> xh(pop * dim);
    HIP_TRY(c, hipMemcpyAsync(kh.data(), d->kb[d->cur], sizeof(unsigned long long) * pop,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(xh.data(), d->xb[d->cur], sizeof(float) * pop * dim,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const unsigned long long k = *std::min_element(kh.begin(), kh.end());
    const size_t i = (size_t)(uint32_t)k;
    const uint32_t fb = (uint32_t)(k >> 32);
    std::memcpy(f, &fb, sizeof fb);
    std::memcpy(x, xh.data() + i * dim, sizeof(float) * dim);
    return MCL_OK;
}

int mcl_de_population(void* de, float* x, float* f)
{
    De* d = static_cast<De*>(de);
    if (!d || !x || !f) return MCL_EINVAL;
    Ctx* c = d->m->ctx;
    HIP_TRY(c, hipMemcpyAsync(x, d->xb[d->cur], sizeof(float) * (size_t)d->a.pop * d->a.dim,
                              hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(f, d->a.f, sizeof(float) * (size_t)d->a.pop,
                              hipMemcpyDeviceToDevice, c->stream));
    return MCL_OK;
}

}  // extern "C"
