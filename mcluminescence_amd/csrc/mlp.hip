// mlp.hip — surrogate network kernels for gfx950: a BF16 matrix-core GEMM with
// fused epilogues (bias, SiLU, loss, SiLU derivative, split-K weight gradient),
// input packing, and the AdamW optimizer step.
//
// Every product is C = A B^T with both operands K-contiguous ("TN"): forward
// X [b][k] . W [n][k]^T, input gradient dZ [b][n] . (W^T) [k][n]^T and weight
// gradient dZ^T [n][b] . X^T [k][b]^T.  The epilogues write the transposed copies
// the next product needs, and the optimizer writes W and W^T, so no kernel ever
// reads a strided operand.
//
// Matrix-core layout of v_mfma_f32_32x32x16_bf16 (wave64): lane l supplies
// A[l % 32][8 (l / 32) .. +7] and B^T[l % 32][8 (l / 32) .. +7]; accumulator
// register r of lane l holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#include <algorithm>

#include "mlp_internal.h"

namespace mcl {
namespace mlp {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kBK = 64;    // K per LDS stage
constexpr int kPad = 8;    // LDS row padding [bf16]: 80-byte rows, conflict-free 16-byte reads


__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup -> output tile.  Workgroups are dispatched round-robin over the 8
// XCDs; renumbering puts consecutive tiles (which share A rows) on one XCD's L2.
__device__ __forceinline__ int xcd_tile(int pid, int total)
{
    if (total % 8 != 0) return pid;
    return (pid % 8) * (total / 8) + pid / 8;
}

// NKT = 0: two LDS stages of kBK, the next fetched while the current is
// multiplied.  NKT > 0 (K = NKT kBK): the whole K extent of the tile is fetched at
// once into NKT LDS stages, one barrier, then multiplied; for products too small
// to fill the device, where each pipelined stage would wait a global-memory round
// trip for a few matrix-core instructions.
template <int BM, int BN, int EPI, int NKT = 0>
__global__ __launch_bounds__(kThreads) void gemm_kernel(GemmArgs g)
{
    constexpr int TM = BM / 2, TN = BN / 2;       // wave tile (2 x 2 waves)
    constexpr int FM = TM / 32, FN = TN / 32;     // 32x32 matrix-core tiles per wave
    constexpr int LDR = kBK + kPad;
    constexpr int CPR = kBK / 8;                  // 16-byte chunks per tile row
    constexpr int CA = BM * CPR / kThreads, CB = BN * CPR / kThreads;
    static_assert(CA >= 1 && CB >= 1, "tile too small for the loader");
    constexpr int NBUF = NKT > 0 ? NKT : 2;
    __shared__ __align__(16) bf16 As[NBUF][BM * LDR];
    __shared__ __align__(16) bf16 Bs[NBUF][BN * LDR];
    __shared__ float red[kThreads / 64];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int tiles_n = g.N / BN;
    const int pid = xcd_tile(blockIdx.x, gridDim.x);
    const int m0 = (pid / tiles_n) * BM, n0 = (pid % tiles_n) * BN;
    const int64_t kbeg = (int64_t)blockIdx.z * g.K;

    const bf16* Ag = g.A + (int64_t)m0 * g.lda + kbeg;
    const bf16* Bg = g.B + (int64_t)n0 * g.ldb + kbeg;
    uint4 ra[CA], rb[CB];
    f32x16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x16{};

    // global -> registers of stage kt (each thread: CA + CB 16-byte chunks)
#define MCL_GLOAD(kt)                                                                              \
    _Pragma("unroll") for (int j = 0; j < CA; ++j) {                                              \
        const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;                               \
        ra[j] = *reinterpret_cast<const uint4*>(Ag + (int64_t)r * g.lda + (kt) * kBK + q * 8);    \
    }                                                                                             \
    _Pragma("unroll") for (int j = 0; j < CB; ++j) {                                              \
        const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;                               \
        rb[j] = *reinterpret_cast<const uint4*>(Bg + (int64_t)r * g.ldb + (kt) * kBK + q * 8);    \
    }
    // registers -> LDS buffer buf
#define MCL_LSTORE(buf)                                                                            \
    _Pragma("unroll") for (int j = 0; j < CA; ++j) {                                              \
        const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;                               \
        *reinterpret_cast<uint4*>(&As[buf][r * LDR + q * 8]) = ra[j];                             \
    }                                                                                             \
    _Pragma("unroll") for (int j = 0; j < CB; ++j) {                                              \
        const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;                               \
        *reinterpret_cast<uint4*>(&Bs[buf][r * LDR + q * 8]) = rb[j];                             \
    }
    // matrix-core products of the stage held in LDS buffer buf
#define MCL_COMPUTE(buf)                                                                           \
    _Pragma("unroll") for (int kk = 0; kk < kBK / 16; ++kk) {                                     \
        bf16x8 a[FM], b[FN];                                                                      \
        const int ko = kk * 16 + (lane >> 5) * 8;                                                 \
        _Pragma("unroll") for (int i = 0; i < FM; ++i) a[i] =                                     \
            *reinterpret_cast<const bf16x8*>(&As[buf][(wm * TM + i * 32 + (lane & 31)) * LDR + ko]); \
        _Pragma("unroll") for (int j = 0; j < FN; ++j) b[j] =                                     \
            *reinterpret_cast<const bf16x8*>(&Bs[buf][(wn * TN + j * 32 + (lane & 31)) * LDR + ko]); \
        _Pragma("unroll") for (int i = 0; i < FM; ++i)                                            \
            _Pragma("unroll") for (int j = 0; j < FN; ++j)                                        \
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0); \
    }

    if constexpr (NKT > 0) {
        // every stage's global loads in flight together, then one barrier
        uint4 sa[NKT][CA], sb[NKT][CB];
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt) {
#pragma unroll
            for (int j = 0; j < CA; ++j) {
                const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;
                sa[kt][j] = *reinterpret_cast<const uint4*>(Ag + (int64_t)r * g.lda + kt * kBK + q * 8);
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;
                sb[kt][j] = *reinterpret_cast<const uint4*>(Bg + (int64_t)r * g.ldb + kt * kBK + q * 8);
            }
        }
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt) {
#pragma unroll
            for (int j = 0; j < CA; ++j) {
                const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;
                *reinterpret_cast<uint4*>(&As[kt][r * LDR + q * 8]) = sa[kt][j];
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const int c = tid + j * kThreads, r = c / CPR, q = c % CPR;
                *reinterpret_cast<uint4*>(&Bs[kt][r * LDR + q * 8]) = sb[kt][j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kt = 0; kt < NKT; ++kt) {
            MCL_COMPUTE(kt)
        }
    } else {
        // Two LDS buffers: stage kt+1 is fetched from global memory while stage kt
        // is multiplied, then written to the other buffer; one barrier per stage.
        const int nk = g.K / kBK;
        MCL_GLOAD(0)
        MCL_LSTORE(0)
        __syncthreads();
        for (int kt = 0; kt + 1 < nk; ++kt) {
            MCL_GLOAD(kt + 1)
            MCL_COMPUTE(kt & 1)
            MCL_LSTORE((kt + 1) & 1)
            __syncthreads();
        }
        MCL_COMPUTE((nk - 1) & 1)
    }
#undef MCL_GLOAD
#undef MCL_LSTORE
#undef MCL_COMPUTE

    // ---- epilogue ---------------------------------------------------------------
    float se = 0.0f;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * TN + j * 32 + (lane & 31);
            const int mb = m0 + wm * TM + i * 32 + (lane >> 5) * 4;
            float bn = 0.0f;
            if constexpr (EPI == kHiddenTrain || EPI == kHiddenInfer || EPI == kOutInfer ||
                          EPI == kOutTrain)
                bn = g.bias[n];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                bf16x4 tv;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mb + 8 * q + r;
                    const int64_t o = (int64_t)m * g.N + n;
                    const float c = acc[i][j][4 * q + r];
                    if constexpr (EPI == kHiddenTrain) {
                        const float z = c + bn;
                        const bf16 a = (bf16)(z * sigmoid(z));
                        g.out_z[o] = (bf16)z;
                        g.out_a[o] = a;
                        tv[r] = a;
                    } else if constexpr (EPI == kHiddenInfer) {
                        const float z = c + bn;
                        g.out_a[o] = (bf16)(z * sigmoid(z));
                    } else if constexpr (EPI == kOutInfer) {
                        g.out_f[o] = c + bn;
                    } else if constexpr (EPI == kOutTrain) {
                        const float y = c + bn;
                        const int64_t trow = g.t_idx ? (int64_t)g.t_idx[m] : (int64_t)m;
                        const float d = y - g.target[trow * g.ld_target + n];
                        se += d * d;
                        const bf16 gd = (bf16)(d * g.grad_scale);
                        g.out_a[o] = gd;
                        tv[r] = gd;
                        if (g.out_f) g.out_f[o] = y;
                    } else if constexpr (EPI == kDxSilu) {
                        const float z = (float)g.z_in[o];
                        const float s = sigmoid(z);
                        const bf16 dz = (bf16)(c * (s * (1.0f + z * (1.0f - s))));
                        if (g.out_a) g.out_a[o] = dz;
                        tv[r] = dz;
                    } else {  // kDwPartial
                        g.out_f[(int64_t)blockIdx.z * g.M * g.N + o] = c;
                    }
                }
                if constexpr (EPI == kHiddenTrain || EPI == kOutTrain || EPI == kDxSilu) {
                    if (g.out_t)
                        *reinterpret_cast<bf16x4*>(g.out_t + (int64_t)n * g.ldt + mb + 8 * q) = tv;
                }
            }
        }
    }
    if constexpr (EPI == kOutTrain) {
        se = wave_sum(se);
        if (lane == 0) red[wave] = se;
        __syncthreads();
        if (tid == 0) g.loss_part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

__global__ __launch_bounds__(kThreads) void pack_input_kernel(const float* x, const int32_t* idx,
                                                              int n_in, int batch, int kp,
                                                              bf16* X, bf16* xt)
{
    const int64_t total = (int64_t)batch * kp;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kThreads) {
        // X row-major [b][k]
        {
            const int b = (int)(e / kp), k = (int)(e % kp);
            const int64_t row = idx ? (int64_t)idx[b] : (int64_t)b;
            X[e] = (bf16)(k < n_in ? x[row * n_in + k] : 0.0f);
        }
        if (xt) {  // transposed [k][b]
            const int k = (int)(e / batch), b = (int)(e % batch);
            const int64_t row = idx ? (int64_t)idx[b] : (int64_t)b;
            xt[e] = (bf16)(k < n_in ? x[row * n_in + k] : 0.0f);
        }
    }
}

__global__ __launch_bounds__(kThreads) void finish_loss_kernel(const float* part, int n, float scale,
                                                               float* loss)
{
    __shared__ float red[kThreads];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += kThreads) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] * scale;
}

__device__ __forceinline__ void adamw(float& w, float& m, float& v, float gr, const OptArgs& a)
{
    w = w * (1.0f - a.lr * a.wd);
    m = a.beta1 * m + (1.0f - a.beta1) * gr;
    v = a.beta2 * v + (1.0f - a.beta2) * (gr * gr);
    w = w - a.step_size * (m / (sqrtf(v) / a.sqrt_bc2 + a.eps));
}

// One workgroup per 32x32 weight tile (W and W^T written through an LDS
// transpose), then one per four biases (one wave each).
__global__ __launch_bounds__(kThreads) void optimizer_kernel(OptArgs a)
{
    __shared__ bf16 tile[32][32 + 2];
    const int blk = blockIdx.x;
    int l = 0;
    bool is_bias = false;
    // layers are laid out as [all weight tiles of every layer][all bias groups]
    for (int i = 0; i < a.n_layers; ++i) {
        const OptLayer& L = a.L[i];
        if (blk >= L.tile0 && blk < L.tile0 + opt_tiles(L.out, L.kp)) { l = i; break; }
        if (blk >= L.bias0 && blk < L.bias0 + opt_bias_groups(L.out)) { l = i; is_bias = true; break; }
    }
    const OptLayer& L = a.L[l];
    const int tid = threadIdx.x;
    if (!is_bias) {
        const int t = blk - L.tile0, tk = L.kp / 32;
        const int tn0 = (t / tk) * 32, tk0 = (t % tk) * 32;
        const int r = tid / 8, c4 = (tid % 8) * 4;
        const int n = tn0 + r, k = tk0 + c4;
        const int64_t o = (int64_t)n * L.kp + k;
        if (a.mode != kOptRefresh) {
            float4 gsum = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int z = 0; z < L.splits; ++z) {
                const float4 p = *reinterpret_cast<const float4*>(L.part + (int64_t)z * L.out * L.kp + o);
                gsum.x += p.x; gsum.y += p.y; gsum.z += p.z; gsum.w += p.w;
            }
            const float gv[4] = {gsum.x, gsum.y, gsum.z, gsum.w};
            if (a.mode == kOptExportGrad) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (k + i < L.k_in) a.flat[L.flat_w + (int64_t)n * L.k_in + k + i] = gv[i];
                return;
            }
            float4 w = *reinterpret_cast<float4*>(L.w + o);
            float4 m = *reinterpret_cast<float4*>(L.m + o);
            float4 v = *reinterpret_cast<float4*>(L.v + o);
            float* wp = &w.x; float* mp = &m.x; float* vp = &v.x;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (k + i < L.k_in) adamw(wp[i], mp[i], vp[i], gv[i], a);
            *reinterpret_cast<float4*>(L.w + o) = w;
            *reinterpret_cast<float4*>(L.m + o) = m;
            *reinterpret_cast<float4*>(L.v + o) = v;
        }
        const float4 w = *reinterpret_cast<const float4*>(L.w + o);
        bf16x4 wb;
        wb[0] = (bf16)w.x; wb[1] = (bf16)w.y; wb[2] = (bf16)w.z; wb[3] = (bf16)w.w;
        *reinterpret_cast<bf16x4*>(L.wb + o) = wb;
#pragma unroll
        for (int i = 0; i < 4; ++i) tile[r][c4 + i] = wb[i];
        __syncthreads();
        // transposed copy: row k of W^T holds column k of W
        const int kr = tid / 8, n4 = (tid % 8) * 4;
        bf16x4 tv;
#pragma unroll
        for (int i = 0; i < 4; ++i) tv[i] = tile[n4 + i][kr];
        *reinterpret_cast<bf16x4*>(L.wt + (int64_t)(tk0 + kr) * L.out + tn0 + n4) = tv;
        return;
    }
    if (a.mode == kOptRefresh) return;
    const int lane = tid & 63;
    const int n = (blk - L.bias0) * 4 + (tid >> 6);
    const bf16* row = L.dzt + (int64_t)n * a.batch;
    float s = 0.0f;
    for (int c = lane; c < a.batch / 8; c += 64) {
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(row + 8 * c);
#pragma unroll
        for (int i = 0; i < 8; ++i) s += (float)v[i];
    }
    s = wave_sum(s);
    if (lane != 0) return;
    if (a.mode == kOptExportGrad) {
        a.flat[L.flat_b + n] = s;
        return;
    }
    adamw(L.bw[n], L.bm[n], L.bv[n], s, a);
}

template <int BM, int BN, int EPI>
hipError_t launch(const GemmArgs& g, int splits, hipStream_t stream)
{
    const int tiles = (g.M / BM) * (g.N / BN);
    hipLaunchKernelGGL((gemm_kernel<BM, BN, EPI>), dim3(tiles, 1, splits), dim3(kThreads), 0,
                       stream, g);
    return hipGetLastError();
}

template <int EPI>
hipError_t launch_tile(const GemmArgs& g, int tile, int splits, hipStream_t stream)
{
    switch (tile) {
    case kTile128x128: return launch<128, 128, EPI>(g, splits, stream);
    case kTile128x64: return launch<128, 64, EPI>(g, splits, stream);
    case kTile64x64: return launch<64, 64, EPI>(g, splits, stream);
    default: return hipErrorInvalidValue;
    }
}

template <int EPI, int NKT>
hipError_t launch_resident(const GemmArgs& g, hipStream_t stream)
{
    const int tiles = (g.M / 64) * (g.N / 64);
    hipLaunchKernelGGL((gemm_kernel<64, 64, EPI, NKT>), dim3(tiles, 1, 1), dim3(kThreads), 0, stream, g);
    return hipGetLastError();
}

template <int EPI>
hipError_t resident_k(const GemmArgs& g, hipStream_t stream)
{
    switch (g.K / kBK) {
    case 1: return launch_resident<EPI, 1>(g, stream);
    case 2: return launch_resident<EPI, 2>(g, stream);
    case 4: return launch_resident<EPI, 4>(g, stream);
    case 8: return launch_resident<EPI, 8>(g, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool gemm_resident_ok(const GemmArgs& g, int tile, int epi, int splits)
{
    const int nk = g.K / kBK;
    return tile == kTile64x64 && splits == 1 && g.K % kBK == 0 &&
           (nk == 1 || nk == 2 || nk == 4 || nk == 8) && (epi == kHiddenInfer || epi == kOutInfer);
}

hipError_t gemm_resident(const GemmArgs& g, int epi, hipStream_t stream)
{
    return epi == kHiddenInfer ? resident_k<kHiddenInfer>(g, stream)
                               : resident_k<kOutInfer>(g, stream);
}

hipError_t gemm(const GemmArgs& g, int tile, int epi, int splits, hipStream_t stream)
{
    switch (epi) {
    case kHiddenTrain: return launch_tile<kHiddenTrain>(g, tile, splits, stream);
    case kHiddenInfer: return launch_tile<kHiddenInfer>(g, tile, splits, stream);
    case kOutInfer: return launch_tile<kOutInfer>(g, tile, splits, stream);
    case kOutTrain: return launch_tile<kOutTrain>(g, tile, splits, stream);
    case kDxSilu: return launch_tile<kDxSilu>(g, tile, splits, stream);
    case kDwPartial: return launch_tile<kDwPartial>(g, tile, splits, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t pack_input(const float* x, const int32_t* idx, int n_in, int batch, int kp, __bf16* X,
                      __bf16* xt, hipStream_t stream)
{
    const int64_t total = (int64_t)batch * kp;
    const int blocks = (int)std::min<int64_t>((total + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks), dim3(kThreads), 0, stream, x, idx, n_in,
                       batch, kp, X, xt);
    return hipGetLastError();
}

hipError_t finish_loss(const float* part, int n, float scale, float* loss, hipStream_t stream)
{
    hipLaunchKernelGGL(finish_loss_kernel, dim3(1), dim3(kThreads), 0, stream, part, n, scale, loss);
    return hipGetLastError();
}

hipError_t optimizer(const OptArgs& a, int n_groups, hipStream_t stream)
{
    hipLaunchKernelGGL(optimizer_kernel, dim3(n_groups), dim3(kThreads), 0, stream, a);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace mcl
