// mlp_fused.hip — the surrogate network's whole inference forward in one kernel.
//
// A workgroup takes 128 rows of the batch through every layer: the rows'
// activations stay in LDS (one BF16 buffer of 128 x lda, rewritten in place by
// each hidden layer after a barrier), the weights come from L2 in fragment order
// (the copy de_pack makes: one 1 KB load per wave per 32-wide output tile and
// 16-wide K step), and only the input rows (FP32) and the output rows (FP32) touch
// HBM.  The layered form writes and re-reads every hidden activation (64 MiB each
// way per hidden layer at batch 65,536).
//
// One workgroup per CU (its LDS is most of the CU's), eight waves; in a layer of
// N outputs wave w takes the 32-wide output tiles w, w + 8 (N = 512: two; N =
// 256: one) for all four 32-row M tiles, so each weight fragment it loads feeds
// four matrix-core products and each activation fragment it reads from LDS feeds
// one or two.  K runs in 16-wide steps in order
// with v_mfma_f32_32x32x16_bf16, the same products in the same order and the
// same epilogue roundings (bias, SiLU, BF16) as the layered GEMM kernels, so the
// outputs are bit-identical to theirs.
#include "mcl_device.h"
#include "mlp_internal.h"

namespace mcl {
namespace mlp {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kRowsF = 128;                  // rows per workgroup
constexpr int kMT = kRowsF / 32;             // 32-row M tiles


// acc[i][j] += rows 32 i .. of act (LDS, stride lda) times the wave's output tile j
// over K = KP, the tile's weights in fragment order at w[j] (the lane's fragment of
// K step 0).  Straight-line code over the K steps (a loop carrying 256
// accumulators through a header makes the register allocator shuttle them), with
// each step's weight fragments fetched PD steps ahead of its products; the
// scheduling barriers keep that order.
template <int FN, int KP, int PD>
__device__ __forceinline__ void fused_products(const bf16* act, int lda, const bf16* const (&w)[FN],
                                               f32x16 (&acc)[kMT][FN])
{
    constexpr int KS = KP / 16, NB = PD + 1;
    const int lane = (int)(threadIdx.x & 63u);
    const bf16* ar = act + (lane & 31) * lda + 8 * (lane >> 5);
    bf16x8 wb[NB][FN];
#pragma unroll
    for (int s = 0; s < PD && s < KS; ++s)
#pragma unroll
        for (int j = 0; j < FN; ++j) wb[s % NB][j] = *reinterpret_cast<const bf16x8*>(w[j] + s * 512);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s + PD < KS) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
                wb[(s + PD) % NB][j] = *reinterpret_cast<const bf16x8*>(w[j] + (s + PD) * 512);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < kMT; ++i) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(ar + i * 32 * lda + 16 * s);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, wb[s % NB][j], acc[i][j], 0, 0, 0);
        }
    }
}

// One layer for the wave's FN output tiles (wave + 4 j): products, then (after
// every wave has read the layer's input) the epilogue: hidden layers write
// bf16(SiLU(z)) into the activation buffer, the last layer writes FP32 rows
// [0, valid) to y (valid is a multiple of 64, so the test is per M tile).
template <int NW, int PD, int FN, int KP, bool LAST>
__device__ __forceinline__ void fused_layer(bf16* act, int lda, const DeLayer& L, float* y,
                                            int64_t row0, int valid, int n_out)
{
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    const bf16* w[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j)
        w[j] = L.w + ((int64_t)(wave + NW * j) * (KP / 16) * 64 + lane) * 8;
    f32x16 acc[kMT][FN];
#pragma unroll
    for (int i = 0; i < kMT; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x16{};
    fused_products<FN, KP, PD>(act, lda, w, acc);
    __syncthreads();                               // the layer's input is read
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int n = 32 * (wave + NW * j) + (lane & 31);
        const float bn = L.b[n];
#pragma unroll
        for (int i = 0; i < kMT; ++i) {
            if (LAST && 32 * i >= valid) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 32 * i + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                const float z = acc[i][j][r] + bn;
                if constexpr (LAST)
                    y[(row0 + m) * n_out + n] = z;
                else
                    act[m * lda + n] = (bf16)(z * sigmoid(z));
            }
        }
    }
    __syncthreads();                               // the layer's output is written
}

template <int NW, int PD>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4)))
void infer_fused_kernel(const float* x, int B, int n_in, DeNet net, float* y, int n_out)
{
    extern __shared__ __align__(16) unsigned char smem_f[];
    bf16* act = reinterpret_cast<bf16*>(smem_f);
    const int lda = net.lda;
    const int64_t row0 = (int64_t)blockIdx.x * kRowsF;
    // input rows: FP32 -> BF16, zero-padded to the first layer's K (and beyond B)
    const int kp0 = net.L[0].kp;
    for (int e = (int)threadIdx.x; e < kRowsF * kp0; e += 64 * NW) {
        const int r = e / kp0, k = e % kp0;
        const int64_t row = row0 + r;
        act[r * lda + k] = (bf16)(k < n_in && row < B ? x[row * n_in + k] : 0.0f);
    }
    __syncthreads();
    // the shapes the host admits: hidden layers 512 wide (the first with 64 padded
    // inputs), the output layer 256 or 512 wide
    constexpr int FH = 16 / NW;                  // tiles a wave in a 512-wide layer
    const int valid = (int)(B - row0 < kRowsF ? B - row0 : kRowsF);
    fused_layer<NW, PD, FH, 64, false>(act, lda, net.L[0], y, row0, valid, n_out);
    for (int l = 1; l + 1 < net.n_layers; ++l)
        fused_layer<NW, PD, FH, 512, false>(act, lda, net.L[l], y, row0, valid, n_out);
    if (net.L[net.n_layers - 1].out == 256)
        fused_layer<NW, PD, FH / 2, 512, true>(act, lda, net.L[net.n_layers - 1], y, row0, valid,
                                               n_out);
    else
        fused_layer<NW, PD, FH, 512, true>(act, lda, net.L[net.n_layers - 1], y, row0, valid, n_out);
}

template <int NW, int PD>
const void* fused_fn()
{
    return reinterpret_cast<const void*>(infer_fused_kernel<NW, PD>);
}

}  // namespace

int infer_fused_lds(const DeNet& net) { return (int)(kRowsF * net.lda * sizeof(bf16)); }

bool infer_fused_ok(const DeNet& net)
{
    if (net.n_layers < 2) return false;
    for (int l = 0; l + 1 < net.n_layers; ++l)
        if (net.L[l].out != 512 || net.L[l].kp != (l == 0 ? 64 : 512)) return false;
    const DeLayer& o = net.L[net.n_layers - 1];
    return (o.out == 256 || o.out == 512) && o.kp == 512;
}

// Eight waves (two per SIMD: one waits on a barrier or its epilogue while the
// other multiplies), weights fetched two K steps ahead.  Measured at batch 65,536:
// 0.154 ms; four waves (one per SIMD, 4 x 4 tiles each) 0.177 ms; eight waves
// fetching four steps ahead 0.164 ms (it spills).
constexpr int kWavesF = 8, kAheadF = 2;

hipError_t infer_fused_prepare(int lds_bytes)
{
    return hipFuncSetAttribute(fused_fn<kWavesF, kAheadF>(),
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

hipError_t infer_fused(const float* x, int B, int n_in, const DeNet& net, float* y, int n_out,
                       int lds_bytes, hipStream_t stream)
{
    hipLaunchKernelGGL((infer_fused_kernel<kWavesF, kAheadF>), dim3((unsigned)((B + kRowsF - 1) / kRowsF)),
                       dim3(64 * kWavesF), (unsigned)lds_bytes, stream, x, B, n_in, net, y, n_out);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace mcl
