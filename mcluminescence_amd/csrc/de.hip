// de.hip — kernels of the network-scored differential evolution: the population
// of the surrogate-assisted fit lives in device memory, and a generation runs
// with no host round trip.
//
// Two forms of a generation, with identical results:
//   - fused (de_generation_kernel): one kernel.  A workgroup takes 32 members:
//     their trials, the whole network forward with the activations in LDS and
//     the weights read from L2 straight into matrix-core registers, the
//     objectives and the selection; 16 waves, one 32-wide output tile each, so
//     for networks whose layers are at most 512 wide.  The weights are read from a copy in fragment order (de_pack_kernel,
//     made when the search is created), so that one load instruction of a wave
//     reads 1 KB of contiguous memory instead of 32 bytes of 32 rows.
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
constexpr int kRows = 32;                    // members per workgroup of the fused kernel
constexpr int kFusedThreads = 1024;          // its 16 waves: one 32-wide output tile each

__device__ __forceinline__ float unit_f(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 0x1p-24f; }

// uniform integer in [0, n) from one 32-bit word (n < 2^31)
__device__ __forceinline__ uint32_t below(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * n) >> 32);
}

// the hidden layers' activation, as the network's GEMM epilogue computes it

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
    __syncthreads();
    k = wmin[0];
    for (int w = 1; w < NT / 64; ++w) k = wmin[w] < k ? wmin[w] : k;
    return (int)(uint32_t)k;
}

// The trial of member i (< pop): its searched coordinates into t[0 .. dim) and the
// network's input row into in[16] (trials, fixed inputs, zeros).
__device__ __forceinline__ void propose_member(const DeArgs& a, int i, int best, uint32_t gen,
                                               float* t, float (&in)[16])
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
    const int jr = (int)below(r.w, (uint32_t)a.dim);
    const float* xi = a.x + (int64_t)i * a.dim;
    const float* x0 = a.x + (int64_t)((i + oa) % P) * a.dim;
    const float* x1 = a.x + (int64_t)((i + ob) % P) * a.dim;
    const float* x2 = a.x + (int64_t)((i + oc) % P) * a.dim;
    const float* xb = a.x + (int64_t)best * a.dim;
    const u32x4 c0 = philox10(gen, (uint32_t)i, 1u, 0u, a.key0, a.key1);   // crossover
    const u32x4 c1 = philox10(gen, (uint32_t)i, 2u, 0u, a.key0, a.key1);
    const u32x4 d0 = philox10(gen, (uint32_t)i, 3u, 0u, a.key0, a.key1);   // redraws
    const u32x4 d1 = philox10(gen, (uint32_t)i, 4u, 0u, a.key0, a.key1);
    float tv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        tv[j] = 0.0f;
        if (j >= a.dim) continue;
        const float m = x0[j] + a.F * (xb[j] - x0[j]) + a.F * (x1[j] - x2[j]);
        const uint32_t uc = j < 4 ? pick_word(c0, j) : pick_word(c1, j - 4);
        const bool cross = unit_f(uc) < a.CR || j == jr;
        float v = cross ? m : xi[j];
        if (v < -1.0f || v > 1.0f) {
            const uint32_t ud = j < 4 ? pick_word(d0, j) : pick_word(d1, j - 4);
            v = 2.0f * unit_f(ud) - 1.0f;
        }
        tv[j] = v;
        t[j] = v;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
        in[k] = k < a.dim ? tv[k & 7] : (k < a.n_in ? a.fixed[k - a.dim] : 0.0f);
}

// Objective of one network output row y [n_out], reduced over the wave (every
// lane returns it): the mean squared difference of the de-normalised curve and
// the target, over the target's squared peak.
__device__ __forceinline__ float member_objective(const DeArgs& a, const float* y, int lane)
{
    float s = 0.0f;
    for (int b = lane; b < a.n_out; b += 64) {
        const float d = expm1_own(y[b] * a.log_span) * a.out_scale - a.target[b];
        s = __builtin_fmaf(d, d, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s / (float)a.n_out * a.inv_peak2;
}

// Greedy selection of member i by one wave: the trial t replaces the member when
// its objective ft is not worse; writes the member's next row and argmin key.
__device__ __forceinline__ void select_member(const DeArgs& a, int i, float ft, const float* t,
                                              int lane)
{
    const float fo = a.f[i];
    const bool take = ft <= fo;
    if (lane < a.dim)
        a.x_next[(int64_t)i * a.dim + lane] = take ? t[lane] : a.x[(int64_t)i * a.dim + lane];
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
    const float ft = member_objective(a, a.Y + (int64_t)i * a.n_out, lane);
    select_member(a, i, ft, a.X + (int64_t)i * a.n_in, lane);
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
// the current one is multiplied.  The fetch past the last chunk re-reads the last
// chunk, so the loop has no branch between a load and the products that wait for
// it.
template <int CH, int NT>
__device__ __forceinline__ void layer_products(const bf16* act, int lda, const bf16* const (&w)[NT],
                                               int kp, f32x16 (&acc)[NT])
{
    const int lane = (int)(threadIdx.x & 63u);
    const bf16* ar = act + (lane & 31) * lda + 8 * (lane >> 5);
    constexpr int KC = 16 * CH;
    bf16x8 p[NT][CH], q[NT][CH];
    static_assert(NT >= 1, "at least one tile");
    auto fetch = [&](bf16x8 (&d)[NT][CH], int k0) {
        k0 = k0 < kp - KC ? k0 : kp - KC;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < CH; ++j)
                d[t][j] = *reinterpret_cast<const bf16x8*>(w[t] + (int64_t)(k0 / 16 + j) * 512);
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
    fetch(p, 0);
    for (int k0 = 0; k0 < kp; k0 += 2 * KC) {
        fetch(q, k0 + KC);
        __builtin_amdgcn_sched_barrier(0);
        multiply(p, k0);
        if (k0 + KC >= kp) break;
        fetch(p, k0 + 2 * KC);
        __builtin_amdgcn_sched_barrier(0);
        multiply(q, k0 + KC);
    }
}

// One whole generation for 32 members per workgroup (see the file comment).
// Dynamic LDS: two activation buffers [32][lda] BF16, the output rows
// [32][n_out + 4] FP32 and the trials [32][8] FP32.
__global__ __launch_bounds__(kFusedThreads) void de_generation_kernel(DeArgs a, DeNet net,
                                                                      uint32_t gen)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lda = net.lda, ys = a.n_out + 4;
    bf16* act0 = reinterpret_cast<bf16*>(smem);
    bf16* act1 = act0 + kRows * lda;
    float* yrow = reinterpret_cast<float*>(act1 + kRows * lda);
    float* trial = yrow + kRows * ys;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = (int)blockIdx.x * kRows;

    const int best = population_argmin<kFusedThreads>(a);   // of the previous generation
    if (tid < kRows) {
        const int i = row0 + tid;
        float in[16];
        if (i < a.pop) {
            propose_member(a, i, best, gen, trial + tid * 8, in);
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) in[k] = 0.0f;   // padding rows
        }
        bf16* r = act0 + tid * lda;
#pragma unroll
        for (int k = 0; k < 16; ++k) r[k] = (bf16)in[k];
        for (int k = 16; k < net.L[0].kp; ++k) r[k] = (bf16)0.0f;
    }
    __syncthreads();

    bf16* cur = act0;
    bf16* nxt = act1;
    for (int l = 0; l < net.n_layers; ++l) {
        const DeLayer L = net.L[l];
        const int t = wave;                       // this wave's output tile
        if (t < L.out / 32) {
            const int KS = L.kp / 16;
            const bf16* const w[1] = {L.w + ((int64_t)t * KS * 64 + lane) * 8};   // fragment order
            f32x16 acc[1] = {f32x16{}};
            // eight K steps per chunk where K allows (more weight bytes in flight)
            if (L.kp % 128 == 0)
                layer_products<4, 1>(cur, lda, w, L.kp, acc);
            else
                layer_products<4, 1>(cur, lda, w, L.kp, acc);
            const bool last = l + 1 == net.n_layers;
            const int n = 32 * t + (lane & 31);
            const float bn = L.b[n];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                if (last) {
                    yrow[m * ys + n] = acc[0][r] + bn;
                } else {
                    const float z = acc[0][r] + bn;
                    nxt[m * lda + n] = (bf16)(z * sigmoid(z));
                }
            }
        }
        __syncthreads();
        bf16* done = cur;
        cur = nxt;
        nxt = done;
    }

    // objectives and selection: wave w takes members 2w and 2w + 1 of the workgroup
    constexpr int kPerWave = kRows / (kFusedThreads / 64);
#pragma unroll
    for (int q = 0; q < kPerWave; ++q) {
        const int r = wave * kPerWave + q, i = row0 + r;
        if (i >= a.pop) continue;
        const float ft = member_objective(a, yrow + r * ys, lane);
        select_member(a, i, ft, trial + r * 8, lane);
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
                 kRows * 8 * sizeof(float));
}

hipError_t de_fused_prepare(int lds_bytes)
{
    return hipFuncSetAttribute(reinterpret_cast<const void*>(de_generation_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

hipError_t de_generation_fused(const DeArgs& a, const DeNet& net, uint32_t gen, int lds_bytes,
                               hipStream_t stream)
{
    hipLaunchKernelGGL(de_generation_kernel, dim3((unsigned)(a.pop_pad / kRows)),
                       dim3(kFusedThreads), (unsigned)lds_bytes, stream, a, net, gen);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace mcl
