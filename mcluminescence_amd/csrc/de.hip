// de.hip — kernels of the network-scored differential evolution: the population
// of the surrogate-assisted fit lives in device memory, and a generation is one
// propose kernel, the network forward (mlp.hip) and one score-and-accept kernel,
// with no host round trip.
//
// The propose kernel writes the trials straight into the network's packed BF16
// input rows (the conversion pack_input makes), so a generation needs no pack pass.
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

constexpr uint32_t kInitGen = 0xFFFFFFFFu;   // counter word 0 of the initial population

__device__ __forceinline__ float unit_f(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 0x1p-24f; }

// uniform integer in [0, n) from one 32-bit word (n < 2^31)
__device__ __forceinline__ uint32_t below(uint32_t w, uint32_t n)
{
    return (uint32_t)(((uint64_t)w * n) >> 32);
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
// the lowest member), reduced by each workgroup: a strided minimum per thread,
// then across the wave by shuffles and across the waves in LDS.
__device__ __forceinline__ int population_argmin(const DeArgs& a)
{
    __shared__ unsigned long long wmin[256 / 64];
    unsigned long long k = ~0ull;
    for (int j = (int)threadIdx.x; j < a.pop; j += (int)blockDim.x) k = a.keys[j] < k ? a.keys[j] : k;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = (unsigned long long)__shfl_xor((long long)k, o, 64);
        k = v < k ? v : k;
    }
    if ((threadIdx.x & 63u) == 0u) wmin[threadIdx.x >> 6] = k;
    __syncthreads();
    k = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) k = wmin[w] < k ? wmin[w] : k;
    return (int)(uint32_t)k;
}

__global__ __launch_bounds__(256) void de_propose_kernel(DeArgs a, uint32_t gen)
{
    const int best = population_argmin(a);          // of the previous generation
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.pop) return;
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
    float* X = a.X + (int64_t)i * a.n_in;
    float in[16];                                    // network inputs 0..15 of the row
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j >= a.dim) break;
        const float m = x0[j] + a.F * (xb[j] - x0[j]) + a.F * (x1[j] - x2[j]);
        const uint32_t uc = j < 4 ? pick_word(c0, j) : pick_word(c1, j - 4);
        const bool cross = unit_f(uc) < a.CR || j == jr;
        float t = cross ? m : xi[j];
        if (t < -1.0f || t > 1.0f) {
            const uint32_t ud = j < 4 ? pick_word(d0, j) : pick_word(d1, j - 4);
            t = 2.0f * unit_f(ud) - 1.0f;
        }
        X[j] = t;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
        in[k] = k < a.dim ? X[k] : (k < a.n_in ? a.fixed[k - a.dim] : 0.0f);
    // the whole packed row, as pack_input writes it (trials, fixed inputs, zeros),
    // eight BF16 values per 16-byte store
    uint4* X0 = reinterpret_cast<uint4*>(a.X0 + (int64_t)i * a.kp0);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        uint32_t w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const __bf16 lo = (__bf16)in[8 * c + 2 * t], hi = (__bf16)in[8 * c + 2 * t + 1];
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
    const float* y = a.Y + (int64_t)i * a.n_out;
    float s = 0.0f;
    for (int b = lane; b < a.n_out; b += 64) {
        const float d = expm1_own(y[b] * a.log_span) * a.out_scale - a.target[b];
        s = __builtin_fmaf(d, d, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const float ft = s / (float)a.n_out * a.inv_peak2;
    const float fo = a.f[i];
    const bool take = ft <= fo;
    if (take && lane < a.dim) a.x[(int64_t)i * a.dim + lane] = a.X[(int64_t)i * a.n_in + lane];
    if (lane == 0) {
        const float fn = take ? ft : fo;
        if (take) a.f[i] = ft;
        a.keys[i] = ((unsigned long long)__float_as_uint(fn) << 32) | (unsigned long long)i;
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

}  // namespace mlp
}  // namespace mcl
