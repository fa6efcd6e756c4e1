// de.hip — kernels of the network-scored differential evolution: the population
// of the surrogate-assisted fit lives in device memory, and a generation runs
// with no host round trip.
//
// Two forms of a generation, with identical results:
//   - fused (de_generation_kernel): one kernel.  A workgroup takes 16 members:
//     their trials, the whole network forward with the activations in LDS and
//     the weights read from L2 straight into matrix-core registers, the
//     objectives and the selection; 16 waves, one 32-wide output tile each, so
//     for networks whose layers are at most 512 wide.  The matrix-core tiles are
//     32 rows: the members' 16 and 16 zero rows, whose outputs no epilogue
//     computes, so that each workgroup's share of the SiLU epilogues and
//     objectives is half of what 32 members would give, over twice the CUs (every
//     CU streams the whole weight set either way).  The weights are read from a
//     copy in fragment order (de_pack_kernel, made when the search is created),
//     so that one load instruction of a wave reads 1 KB of contiguous memory
//     instead of 32 bytes of 32 rows.
//   - layered: a propose kernel that writes the trials into the network's packed
//     BF16 input rows, the network's own GEMM kernels (mlp.hip), and a
//     score-and-accept kernel.
// Both run the same matrix-core products in the same K order with the same
// roundings (BF16 activations, the same SiLU), and the same objective sum.
//
// The population is double-buffered: a generation reads members and argmin keys
// from (x, keys) and writes the selected members and their keys to (x_next,
// keys_next), so a trial never reads a partner that another workgroup of the same
// generation has already replaced.
//
// DE/rand-to-best/1/bin, the step of scipy's differential_evolution that the
// reference's fit calls (src/class/optimizer.py:98-111):
//   mutant  = x_a + F (x_best - x_a) + F (x_b - x_c), a, b, c distinct and != i
//   trial_j = mutant_j if U_j < CR or j == j_rand, else x_ij
//   a coordinate outside [-1, 1] is redrawn uniformly
//   the trial replaces x_i when its objective is not worse.
// Draws: Philox4x32-10 with key = seed, counter = (generation, member, block, 0).
#include "mcl_device.h"
#include "mlp_internal.h"

namespace mcl {
namespace mlp {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kInitGen = 0xFFFFFFFFu;   // counter word 0 of the initial population
constexpr int kRows = 32;                    // rows of the fused kernel's matrix-core tiles
constexpr int kMembers = 16;                 // members per workgroup: tile rows 0-15 (16-31 are
                                             // zero padding rows, which no epilogue writes)
constexpr int kFusedThreads = 1024;          // its 16 waves: one 32-wide output tile each

__device__ __forceinline__ float unit_f(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 0x1p-24f; }

// uniform integer in [0, n) from one 32-bit word (n < 2^31)
__device__ __forceinline__ uint32_t below(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * n) >> 32);
}

// Workgroup barrier that orders LDS only: the wave's global loads stay in flight
// across it (a full barrier waits for every outstanding load, since the vector
// memory counter counts loads and stores alike).  Every exchange between the
// waves of these kernels goes through LDS; the global data they read (population,
// keys, weights) does not change during a generation.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__global__ __launch_bounds__(256) void de_init_kernel(DeArgs a)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.pop_pad) return;
    float* X = a.X + (int64_t)i * a.n_in;
    if (i >= a.pop) {
        for (int j = 0; j < a.n_in; ++j) X[j] = 0.0f;   // padding rows of the forward batch
        return;
    }
    const u32x4 w0 = philox10(kInitGen, (uint32_t)i, 0u, 0u, a.key0, a.key1);
    const u32x4 w1 = philox10(kInitGen, (uint32_t)i, 1u, 0u, a.key0, a.key1);
    for (int j = 0; j < a.dim; ++j) {
        const uint32_t w = j < 4 ? pick_word(w0, j) : pick_word(w1, j - 4);
        const float v = 2.0f * unit_f(w) - 1.0f;
        X[j] = v;
        a.x[(int64_t)i * a.dim + j] = v;
    }
    for (int k = 0; a.dim + k < a.n_in; ++k) X[a.dim + k] = a.fixed[k];
    a.f[i] = __builtin_inff();
}

// Population argmin from the members' keys (objective bits << 32 | member; the
// objectives are >= 0, so their float bits order as the values, and ties go to
// the lowest member), reduced by each workgroup of NT threads: a strided minimum
// per thread, then across the wave by shuffles and across the waves in LDS.
template <int NT>
__device__ __forceinline__ int population_argmin(const DeArgs& a)
{
    __shared__ unsigned long long wmin[NT / 64];
    unsigned long long k = ~0ull;
    for (int j = (int)threadIdx.x; j < a.pop; j += NT) k = a.keys[j] < k ? a.keys[j] : k;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = (unsigned long long)__shfl_xor((long long)k, o, 64);
        k = v < k ? v : k;
    }
    if ((threadIdx.x & 63u) == 0u) wmin[threadIdx.x >> 6] = k;
    lds_barrier();
    k = wmin[0];
    for (int w = 1; w < NT / 64; ++w) k = wmin[w] < k ? wmin[w] : k;
    return (int)(uint32_t)k;
}

// The trial of member i (< pop) in two halves, so that its draws and partner
// rows can be fetched before the population argmin is known: trial_fetch takes
// the draws and the rows of the member and its three partners, trial_finish the
// best member's row and the mutant, crossover and redraws.  Together they are
// the one-step computation the layered form's propose kernel does.
struct Trial {
    u32x4 c0, c1, d0, d1;       // crossover and redraw words
    int jr;                     // the coordinate that always crosses over
    float xi[8], x0[8], x1[8], x2[8];
};

__device__ __forceinline__ void trial_fetch(const DeArgs& a, int i, uint32_t gen, Trial& T)
{
    const uint32_t P = (uint32_t)a.pop;
    const u32x4 r = philox10(gen, (uint32_t)i, 0u, 0u, a.key0, a.key1);
    // three distinct offsets in [1, P): partners differ from i and from one another
    const uint32_t oa = 1u + below(r.x, P - 1u);
    uint32_t ob = 1u + below(r.y, P - 2u);
    ob += ob >= oa ? 1u : 0u;
    uint32_t oc = 1u + below(r.z, P - 3u);
    const uint32_t lo = oa < ob ? oa : ob, hi = oa < ob ? ob : oa;
    oc += oc >= lo ? 1u : 0u;
    oc += oc >= hi ? 1u : 0u;
    T.jr = (int)below(r.w, (uint32_t)a.dim);
    const float* xi = a.x + (int64_t)i * a.dim;
    const float* x0 = a.x + (int64_t)((i + oa) % P) * a.dim;
    const float* x1 = a.x + (int64_t)((i + ob) % P) * a.dim;
    const float* x2 = a.x + (int64_t)((i + oc) % P) * a.dim;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool in = j < a.dim;
        T.xi[j] = in ? xi[j] : 0.0f;
        T.x0[j] = in ? x0[j] : 0.0f;
        T.x1[j] = in ? x1[j] : 0.0f;
        T.x2[j] = in ? x2[j] : 0.0f;
    }
    T.c0 = philox10(gen, (uint32_t)i, 1u, 0u, a.key0, a.key1);   // crossover
    T.c1 = philox10(gen, (uint32_t)i, 2u, 0u, a.key0, a.key1);
    T.d0 = philox10(gen, (uint32_t)i, 3u, 0u, a.key0, a.key1);   // redraws
    T.d1 = philox10(gen, (uint32_t)i, 4u, 0u, a.key0, a.key1);
}

// Its searched coordinates into t[0 .. dim) and the network's input row into
// in[16] (trials, fixed inputs, zeros).
__device__ __forceinline__ void trial_finish(const DeArgs& a, const Trial& T, int best, float* t,
                                             float (&in)[16])
{
    const float* xb = a.x + (int64_t)best * a.dim;
    float tv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        tv[j] = 0.0f;
        if (j >= a.dim) continue;
        const float m = T.x0[j] + a.F * (xb[j] - T.x0[j]) + a.F * (T.x1[j] - T.x2[j]);
        const uint32_t uc = j < 4 ? pick_word(T.c0, j) : pick_word(T.c1, j - 4);
        const bool cross = unit_f(uc) < a.CR || j == T.jr;
        float v = cross ? m : T.xi[j];
        if (v < -1.0f || v > 1.0f) {
            const uint32_t ud = j < 4 ? pick_word(T.d0, j) : pick_word(T.d1, j - 4);
            v = 2.0f * unit_f(ud) - 1.0f;
        }
        tv[j] = v;
        t[j] = v;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
        in[k] = k < a.dim ? tv[k & 7] : (k < a.n_in ? a.fixed[k - a.dim] : 0.0f);
}

__device__ __forceinline__ void propose_member(const DeArgs& a, int i, int best, uint32_t gen,
                                               float* t, float (&in)[16])
{
    Trial T;
    trial_fetch(a, i, gen, T);
    trial_finish(a, T, best, t, in);
}

// Objective of one network output row y [n_out], reduced over the wave (every
// lane returns it): the mean squared difference of the de-normalised curve and
// the target (tgt: a.target or a copy of it), over the target's squared peak.
__device__ __forceinline__ float member_objective(const DeArgs& a, const float* y, const float* tgt,
                                                  int lane)
{
    float s = 0.0f;
    for (int b = lane; b < a.n_out; b += 64) {
        const float d = expm1_own(y[b] * a.log_span) * a.out_scale - tgt[b];
        s = __builtin_fmaf(d, d, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s / (float)a.n_out * a.inv_peak2;
}

// Greedy selection of member i by one wave: the trial t replaces the member's
// row xi when its objective ft is not worse than fo (its current one, a.f[i]);
// writes the member's next row and argmin key.
__device__ __forceinline__ void select_member(const DeArgs& a, int i, float ft, float fo,
                                              const float* t, const float* xi, int lane)
{
    const bool take = ft <= fo;
    if (lane < a.dim) a.x_next[(int64_t)i * a.dim + lane] = take ? t[lane] : xi[lane];
    if (lane == 0) {
        const float fn = take ? ft : fo;
        if (take) a.f[i] = ft;
        a.keys_next[i] = ((unsigned long long)__float_as_uint(fn) << 32) | (unsigned long long)i;
    }
}

__global__ __launch_bounds__(256) void de_propose_kernel(DeArgs a, uint32_t gen)
{
    const int best = population_argmin<256>(a);     // of the previous generation
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.pop) return;
    float in[16];
    propose_member(a, i, best, gen, a.X + (int64_t)i * a.n_in, in);
    // the whole packed row, as pack_input writes it (trials, fixed inputs, zeros),
    // eight BF16 values per 16-byte store
    uint4* X0 = reinterpret_cast<uint4*>(a.X0 + (int64_t)i * a.kp0);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        uint32_t w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16 lo = (bf16)in[8 * c + 2 * t], hi = (bf16)in[8 * c + 2 * t + 1];
            w[t] = (uint32_t)__builtin_bit_cast(unsigned short, lo) |
                   ((uint32_t)__builtin_bit_cast(unsigned short, hi) << 16);
        }
        X0[c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    for (int c = 2; c < a.kp0 / 8; ++c) X0[c] = make_uint4(0u, 0u, 0u, 0u);
}

// One wave per member: objective of its trial (the network output row), greedy
// selection, and the member's argmin key.
__global__ __launch_bounds__(256) void de_score_kernel(DeArgs a)
{
    const int lane = (int)(threadIdx.x & 63u);
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (i >= a.pop) return;
    const float ft = member_objective(a, a.Y + (int64_t)i * a.n_out, a.target, lane);
    select_member(a, i, ft, a.f[i], a.X + (int64_t)i * a.n_in, a.x + (int64_t)i * a.dim, lane);
}

// Weights in fragment order: for output tile t (32 rows), K step s (16 columns)
// and lane l, the 8 BF16 values W[32 t + l % 32][16 s + 8 (l / 32) .. +7] that lane
// l supplies to the 32x32x16 product, at ((t KS + s) 64 + l) 8, KS = kp / 16.
__global__ __launch_bounds__(256) void de_pack_kernel(const bf16* W, int out, int kp, bf16* Wf)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // (t, s, l)
    const int KS = kp / 16;
    if (g >= (int64_t)(out / 32) * KS * 64) return;
    const int l = (int)(g % 64), s = (int)((g / 64) % KS), t = (int)(g / (64 * KS));
    *reinterpret_cast<uint4*>(Wf + g * 8) = *reinterpret_cast<const uint4*>(
        W + (int64_t)(32 * t + (l & 31)) * kp + 16 * s + 8 * (l >> 5));
}

// Matrix-core products of one layer for the wave's NT output tiles (w[j] points
// at the lane's fragment of K step 0 of tile j, in fragment order) over the 32
// rows in LDS (act, row stride lda): acc[j][r] = sum_k act[m][k] W[n][k], with m
// and n where the 32x32x16 layout puts them (lane l supplies A[l % 32][8 (l / 32)
// .. +7] and B^T[l % 32][8 (l / 32) .. +7]).  K in steps of 16 in order, CH steps
// per chunk (kp a multiple of 16 CH); the next chunk's weights are fetched while
// the current one is multiplied, alternating between p and q.  p holds the
// layer's first chunk on entry (fetched before the layer's input was ready).  The
// fetch past the last chunk is the next layer's first chunk (wn, when next; else
// the last chunk again), left in flight over the epilogue and the barrier: in p
// after an even number of chunks, in q after an odd one (returned).  The loads
// are unconditional (a uniform choice of address), so the wait before each
// chunk's products counts exactly the loads issued after it.
template <int CH, int NT>
__device__ __forceinline__ bool layer_products(const bf16* act, int lda, const bf16* const (&w)[NT],
                                               int kp, f32x16 (&acc)[NT], bf16x8 (&p)[NT][CH],
                                               bf16x8 (&q)[NT][CH], const bf16* const (&wn)[NT],
                                               bool next)
{
    const int lane = (int)(threadIdx.x & 63u);
    const bf16* ar = act + (lane & 31) * lda + 8 * (lane >> 5);
    constexpr int KC = 16 * CH;
    static_assert(NT >= 1, "at least one tile");
    auto fetch = [&](bf16x8 (&d)[NT][CH], int k0) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16* src = k0 < kp ? w[t] + (int64_t)(k0 / 16) * 512
                                      : (next ? wn[t] : w[t] + (int64_t)((kp - KC) / 16) * 512);
#pragma unroll
            for (int j = 0; j < CH; ++j) d[t][j] = *reinterpret_cast<const bf16x8*>(src + j * 512);
        }
    };
    auto multiply = [&](const bf16x8 (&d)[NT][CH], int k0) {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(ar + k0 + 16 * j);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, d[t][j], acc[t], 0, 0, 0);
        }
    };
    // the scheduling barriers keep each chunk's loads ahead of the previous chunk's
    // products (the scheduler would otherwise sink them next to their own uses)
    for (int k0 = 0; k0 < kp; k0 += 2 * KC) {
        fetch(q, k0 + KC);
        __builtin_amdgcn_sched_barrier(0);
        multiply(p, k0);
        if (k0 + KC >= kp) return true;
        fetch(p, k0 + 2 * KC);
        __builtin_amdgcn_sched_barrier(0);
        multiply(q, k0 + KC);
    }
    return false;
}

// The lane's fragment of K step 0 of output tile t of layer L, in fragment order.
__device__ __forceinline__ const bf16* tile_fragments(const DeLayer& L, int t, int lane)
{
    return L.w + ((int64_t)t * (L.kp / 16) * 64 + lane) * 8;
}

// One whole generation for 32 members per workgroup (see the file comment).
// Dynamic LDS: two activation buffers [32][lda] BF16, the output rows
// [32][n_out + 4] FP32, the trials [32][8] FP32, the target [n_out], the
// members' current objectives [32] and rows [32][8] FP32.
// The fixed network shapes of the unrolled form (the surrogate networks): five
// layers, 64 padded inputs, hidden layers 512 wide, 256 outputs.
constexpr int kFixedLayers = 5, kLdaFixed = 512 + 8, kYsFixed = 256 + 4;
constexpr int kAheadDe = 1;     // chunks in flight ahead of the products (4 KB a wave)

// The fixed shapes as one stream of weight chunks (four K steps, 4 KB per wave):
// chunk 0 is layer 0's only chunk (K = 64), chunks 1 + 8 (l - 1) .. + 7 are layer
// l's (K = 512).  wl[l] is the wave's tile of layer l in fragment order.
constexpr int kFixedChunks = 8, kFixedStream = 1 + kFixedChunks * (kFixedLayers - 1);

__device__ __forceinline__ const bf16* stream_chunk(const bf16* const (&wl)[kFixedLayers], int g)
{
    return g == 0 ? wl[0] : wl[1 + (g - 1) / kFixedChunks] + ((g - 1) % kFixedChunks) * 2048;
}

// Layer L of the fixed shapes (bn: the bias of the lane's output column, loaded at
// the kernel's start), its products unrolled over the chunk stream: chunk
// g sits in buf[g % D] and is fetched D - 1 chunks ahead of its products, across
// the layer boundaries too (the next layer's first chunks are in flight over the
// epilogue and the barrier).  Every buffer index is known at compile time, so
// the ring needs no register copies; within a layer the loads are unconditional,
// so the wait before each chunk's products counts exactly the loads issued after
// it (the output layer's eight idle waves skip it whole).  The same products in
// the same K order as layer_products.
template <int L, int D, int LDA, int YS>
__device__ __forceinline__ void fixed_layer(const bf16* const (&wl)[kFixedLayers], float bn,
                                            bf16x8 (&buf)[D][4], bf16*& cur, bf16*& nxt,
                                            float* yrow, int t, int lane)
{
    constexpr bool kLastL = L == kFixedLayers - 1;
    constexpr int G0 = L == 0 ? 0 : 1 + kFixedChunks * (L - 1), NC = L == 0 ? 1 : kFixedChunks;
    const bool has = !kLastL || t < 8;                // the output layer has 8 tiles
    const int n = 32 * (kLastL ? (t & 7) : t) + (lane & 31);
    const bf16* ar = cur + (lane & 31) * LDA + 8 * (lane >> 5);
    auto fetch = [&](int gf) {
        const bf16* src = stream_chunk(wl, gf);
#pragma unroll
        for (int j = 0; j < 4; ++j) buf[gf % D][j] = *reinterpret_cast<const bf16x8*>(src + j * 512);
    };
    if (has) {
        f32x16 acc = f32x16{};
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int g = G0 + c, gf = g + D - 1;
            // a chunk of this layer goes out before the products of chunk g; one of
            // the next layer after them (its products wait for the barrier anyway),
            // and only by the waves that have a tile there
            const bool cross = gf >= G0 + NC;
            if (!cross && gf < kFixedStream) fetch(gf);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bf16x8 av = *reinterpret_cast<const bf16x8*>(ar + 64 * c + 16 * j);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, buf[g % D][j], acc, 0, 0, 0);
            }
            if (cross && gf < kFixedStream) {
                const bool next_last = gf >= 1 + kFixedChunks * (kFixedLayers - 2);
                if (!next_last || t < 8) fetch(gf);
            }
        }
        // accumulator r holds row 8 (r / 4) + 4 (lane / 32) + r % 4: the members'
        // rows 0-15 are r < 8
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int m = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
            if constexpr (kLastL) {
                yrow[m * YS + n] = acc[r] + bn;
            } else {
                const float z = acc[r] + bn;
                nxt[m * LDA + n] = (bf16)(z * sigmoid(z));
            }
        }
    }
    lds_barrier();
    bf16* done = cur;
    cur = nxt;
    nxt = done;
}

// LDA, YS: the activation and output row strides when fixed at compile time (the
// fixed shapes, laid out by a fully unrolled forward), so that the epilogue's LDS
// addresses are one base and immediate offsets; 0: any shape the fused kernel
// takes, net.lda and n_out + 4 at run time.
template <int LDA, int YS>
__global__ __launch_bounds__(kFusedThreads) void de_generation_kernel(DeArgs a, DeNet net,
                                                                      uint32_t gen)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lda = LDA > 0 ? LDA : net.lda, ys = YS > 0 ? YS : a.n_out + 4;
    bf16* act0 = reinterpret_cast<bf16*>(smem);
    bf16* act1 = act0 + kRows * lda;
    float* yrow = reinterpret_cast<float*>(act1 + kRows * lda);
    float* trial = yrow + kRows * ys;
    float* tgt = trial + kRows * 8;           // [n_out] the target curve
    float* fcur = tgt + a.n_out;              // [kRows] the members' current objectives
    float* xown = fcur + kRows;               // [kRows][8] the members' current rows
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = (int)blockIdx.x * kMembers;
    constexpr int kChunk = 4;                // K steps per weight chunk
    const int t = wave;                      // the wave's output tile in every layer
    __builtin_assume(t < kFusedThreads / 64);   // every wave has a tile of a hidden layer

    // the fixed shapes: the wave's tiles of every layer, and a ring of kAheadDe + 1
    // chunk buffers over the stream whose first chunks go out first of all
    constexpr int D = kAheadDe + 1;
    const bf16* wl[kFixedLayers];
    float bl[kFixedLayers];                  // the lane's output-column bias of every layer
    bf16x8 buf[D][4];
    if constexpr (LDA > 0) {
#pragma unroll
        for (int l = 0; l < kFixedLayers; ++l) {
            const int tl = l + 1 == kFixedLayers ? (t & 7) : t;
            wl[l] = tile_fragments(net.L[l], tl, lane);
            bl[l] = net.L[l].b[32 * tl + (lane & 31)];
        }
#pragma unroll
        for (int g = 0; g < D - 1; ++g) {
            const bf16* src = stream_chunk(wl, g);
#pragma unroll
            for (int j = 0; j < 4; ++j) buf[g][j] = *reinterpret_cast<const bf16x8*>(src + j * 512);
        }
    }

    // the trials' draws and partner rows, the target and the members' current
    // objectives go out before the argmin, so their latencies overlap
    Trial T;
    if (tid < kMembers && row0 + tid < a.pop) trial_fetch(a, row0 + tid, gen, T);
    for (int b = tid; b < a.n_out; b += kFusedThreads) tgt[b] = a.target[b];
    if (tid < kMembers) fcur[tid] = row0 + tid < a.pop ? a.f[row0 + tid] : 0.0f;
    // the padding rows of both activation buffers hold zeros (16-byte stores; the
    // row stride lda is a multiple of 8)
    for (int e = tid; e < (kRows - kMembers) * lda / 8; e += kFusedThreads) {
        reinterpret_cast<uint4*>(act0 + kMembers * lda)[e] = make_uint4(0u, 0u, 0u, 0u);
        reinterpret_cast<uint4*>(act1 + kMembers * lda)[e] = make_uint4(0u, 0u, 0u, 0u);
    }
    const int best = population_argmin<kFusedThreads>(a);   // of the previous generation
    if (tid < kMembers) {
        const int i = row0 + tid;
        float in[16];
        if (i < a.pop) {
            trial_finish(a, T, best, trial + tid * 8, in);
#pragma unroll
            for (int j = 0; j < 8; ++j) xown[tid * 8 + j] = T.xi[j];
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) in[k] = 0.0f;   // padding rows
        }
        bf16* r = act0 + tid * lda;
#pragma unroll
        for (int k = 0; k < 16; ++k) r[k] = (bf16)in[k];
        for (int k = 16; k < net.L[0].kp; ++k) r[k] = (bf16)0.0f;
    }
    // wave w computes output tile w of every layer that is that wide.  The layer
    // loop is unrolled (layer l's descriptor at a fixed kernel-argument offset:
    // scalar loads, which leave the weight prefetch in flight), and each layer's
    // bias is fetched before its products, so waiting for it does not wait for the
    // next layer's first chunk.
    bf16* cur = act0;
    bf16* nxt = act1;
    if constexpr (LDA > 0) {
        lds_barrier();
        static_assert(kFixedLayers == 5, "the unrolled form is written for five layers");
        fixed_layer<0, D, LDA, YS>(wl, bl[0], buf, cur, nxt, yrow, t, lane);
        fixed_layer<1, D, LDA, YS>(wl, bl[1], buf, cur, nxt, yrow, t, lane);
        fixed_layer<2, D, LDA, YS>(wl, bl[2], buf, cur, nxt, yrow, t, lane);
        fixed_layer<3, D, LDA, YS>(wl, bl[3], buf, cur, nxt, yrow, t, lane);
        fixed_layer<4, D, LDA, YS>(wl, bl[4], buf, cur, nxt, yrow, t, lane);
    } else {
        bf16x8 pre[1][kChunk], alt[1][kChunk];
        if (t < net.L[0].out / 32) {
            const bf16* w0 = tile_fragments(net.L[0], t, lane);
#pragma unroll
            for (int j = 0; j < kChunk; ++j)
                pre[0][j] = *reinterpret_cast<const bf16x8*>(w0 + j * 512);
        }
        lds_barrier();
        bool in_alt = false;          // the layer's first chunk is in alt, not pre
#pragma unroll
        for (int l = 0; l < kMaxLayers; ++l) {
            if (l >= net.n_layers) break;
            const DeLayer L = net.L[l];
            if (t < L.out / 32) {
                if (in_alt) {
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) pre[0][j] = alt[0][j];
                }
                const int n = 32 * t + (lane & 31);
                const float bn = L.b[n];
                const bf16* const w[1] = {tile_fragments(L, t, lane)};
                const int ln = l + 1 < kMaxLayers ? l + 1 : l;
                const bool next = l + 1 < net.n_layers && t < net.L[ln].out / 32;
                const bf16* const wn[1] = {next ? tile_fragments(net.L[ln], t, lane) : w[0]};
                f32x16 acc[1] = {f32x16{}};
                in_alt = layer_products<kChunk, 1>(cur, lda, w, L.kp, acc, pre, alt, wn, next);
                const bool last = l + 1 == net.n_layers;
#pragma unroll
                for (int r = 0; r < 8; ++r) {      // the members' rows 0-15
                    const int m = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                    if (last) {
                        yrow[m * ys + n] = acc[0][r] + bn;
                    } else {
                        const float z = acc[0][r] + bn;
                        nxt[m * lda + n] = (bf16)(z * sigmoid(z));
                    }
                }
            }
            lds_barrier();
            bf16* done = cur;
            cur = nxt;
            nxt = done;
        }
    }

    // objectives and selection: wave w takes member w of the workgroup
    constexpr int kPerWave = kMembers / (kFusedThreads / 64);
#pragma unroll
    for (int q = 0; q < kPerWave; ++q) {
        const int r = wave * kPerWave + q, i = row0 + r;
        if (i >= a.pop) continue;
        const float ft = member_objective(a, yrow + r * ys, tgt, lane);
        select_member(a, i, ft, fcur[r], trial + r * 8, xown + r * 8, lane);
    }
}

}  // namespace

hipError_t de_init(const DeArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL(de_init_kernel, dim3((unsigned)((a.pop_pad + 255) / 256)), dim3(256), 0,
                       stream, a);
    return hipGetLastError();
}

hipError_t de_propose(const DeArgs& a, uint32_t gen, hipStream_t stream)
{
    hipLaunchKernelGGL(de_propose_kernel, dim3((unsigned)((a.pop + 255) / 256)), dim3(256), 0,
                       stream, a, gen);
    return hipGetLastError();
}

hipError_t de_score(const DeArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL(de_score_kernel, dim3((unsigned)((a.pop + 3) / 4)), dim3(256), 0, stream,
                       a);
    return hipGetLastError();
}

hipError_t de_pack(const __bf16* W, int out, int kp, __bf16* Wf, hipStream_t stream)
{
    const int64_t n = (int64_t)(out / 32) * (kp / 16) * 64;
    hipLaunchKernelGGL(de_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, W,
                       out, kp, Wf);
    return hipGetLastError();
}

int de_fused_lds(const DeNet& net, int n_out)
{
    return (int)(2 * kRows * net.lda * sizeof(bf16) + kRows * (n_out + 4) * sizeof(float) +
                 kRows * 8 * sizeof(float) + n_out * sizeof(float) + kRows * sizeof(float) +
                 kRows * 8 * sizeof(float));
}


hipError_t de_fused_prepare(int lds_bytes)
{
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(de_generation_kernel<kLdaFixed, kYsFixed>),
        hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(de_generation_kernel<0, 0>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
    return e;
}

// Whether a network has the fixed shapes of the unrolled form.
bool de_fixed_shapes(const DeNet& net, int n_out)
{
    if (net.n_layers != kFixedLayers || net.lda != kLdaFixed || n_out + 4 != kYsFixed) return false;
    for (int l = 0; l < net.n_layers; ++l) {
        const bool last = l + 1 == net.n_layers;
        if (net.L[l].kp != (l == 0 ? 64 : 512) || net.L[l].out != (last ? 256 : 512)) return false;
    }
    return true;
}

hipError_t de_generation_fused(const DeArgs& a, const DeNet& net, uint32_t gen, int lds_bytes,
                               hipStream_t stream)
{
    const dim3 grid((unsigned)(a.pop_pad / kMembers)), block(kFusedThreads);
    if (de_fixed_shapes(net, a.n_out))
        hipLaunchKernelGGL((de_generation_kernel<kLdaFixed, kYsFixed>), grid, block,
                           (unsigned)lds_bytes, stream, a, net, gen);
    else
        hipLaunchKernelGGL((de_generation_kernel<0, 0>), grid, block, (unsigned)lds_bytes, stream,
                           a, net, gen);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace mcl
