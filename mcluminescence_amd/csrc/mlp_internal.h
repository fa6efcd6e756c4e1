// mlp_internal.h — descriptors shared by the surrogate kernels (mlp.hip) and
// their host side (mlp_abi.hip); not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcl {
namespace mlp {

// The logistic function of the hidden layers' SiLU, z * sigmoid(z), as every
// kernel computes it (GEMM epilogues, the fused forward, the DE generation):
// the hardware reciprocal (1 ulp) of 1 + e^{-z}; one shared definition keeps
// their results bit-identical.
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// What a GEMM's epilogue does with the FP32 accumulator acc[m][n] of
// C = A B^T (A [M][K], B [N][K], both K-contiguous BF16).
enum Epi : int {
    kHiddenTrain = 0,  // z = acc + bias: out_z = z, out_a = silu(z), out_t = silu(z)^T
    kHiddenInfer = 1,  // out_a = silu(acc + bias)
    kOutInfer = 2,     // out_f = acc + bias (FP32)
    kOutTrain = 3,     // y = acc + bias; d = y - target: out_a = d * grad_scale, out_t = its
                       // transpose, loss_part[block] = sum d^2 (out_f = y if given)
    kDxSilu = 4,       // dz = acc * silu'(z_in): out_a = dz (if given), out_t = dz^T
    kDwPartial = 5,    // out_f[split][M][N] = acc over the split's K range
    kEpiCount = 6
};

// Tile shapes (rows x columns of C per 256-thread workgroup).
enum Tile : int { kTile128x128 = 0, kTile128x64 = 1, kTile64x64 = 2, kTileCount = 3 };
inline int tile_m(int t) { return t == kTile64x64 ? 64 : 128; }
inline int tile_n(int t) { return t == kTile128x128 ? 128 : 64; }

struct GemmArgs {
    const __bf16* A;          // [M][lda], this split's K range starts at z * K
    const __bf16* B;          // [N][ldb]
    int64_t lda, ldb;
    int M, N, K;              // K = contraction length of one split (multiple of 64)
    const float* bias;        // [N]
    __bf16* out_a;            // [M][N]
    __bf16* out_z;            // [M][N]
    __bf16* out_t;            // [N][ldt]
    int64_t ldt;
    float* out_f;             // [M][N], or [splits][M][N] partial sums
    const __bf16* z_in;       // [M][N]
    const float* target;      // [rows][ld_target]
    const int32_t* t_idx;     // target row of output row m (null: m)
    int64_t ld_target;
    float grad_scale;
    float* loss_part;         // [workgroups]
};

hipError_t gemm(const GemmArgs& g, int tile, int epi, int splits, hipStream_t stream);
// The whole-K variant for products too small to fill the device: 64 x 64 tiles,
// K = 64, 128, 256 or 512 held in LDS at once, inference epilogues only.
bool gemm_resident_ok(const GemmArgs& g, int tile, int epi, int splits);
hipError_t gemm_resident(const GemmArgs& g, int epi, hipStream_t stream);

// x [rows][n_in] FP32 (row idx[b], or b) -> X [batch][kp] BF16 zero-padded and,
// if xt is given, X^T [kp][batch].
hipError_t pack_input(const float* x, const int32_t* idx, int n_in, int batch, int kp,
                      __bf16* X, __bf16* xt, hipStream_t stream);

// loss = sum(part[0..n)) * scale, summed in a fixed order.
hipError_t finish_loss(const float* part, int n, float scale, float* loss, hipStream_t stream);

constexpr int kMaxLayers = 8;

struct OptLayer {
    float* w; float* m; float* v;     // FP32 master weight and moments [out][kp]
    __bf16* wb;                       // BF16 copy [out][kp]
    __bf16* wt;                       // BF16 transposed copy [kp][out]
    const float* part;                // weight-gradient partial sums [splits][out][kp]
    int splits;
    float* bw; float* bm; float* bv;  // bias master and moments [out]
    const __bf16* dzt;                // dL/dz^T [out][batch] (bias gradient = row sums)
    int out, kp, k_in;                // k_in: unpadded inputs (flat layout)
    int tile0, bias0;                 // first weight-tile / bias workgroup of the layer
    int64_t flat_w, flat_b;           // offsets of W and b in the flat layout
};

enum OptMode : int { kOptAdamW = 0, kOptExportGrad = 1, kOptRefresh = 2 };

struct OptArgs {
    OptLayer L[kMaxLayers];
    int n_layers;
    int mode;
    int batch;
    float* flat;                      // kOptExportGrad destination
    float lr, beta1, beta2, eps, wd;
    float step_size;                  // lr / (1 - beta1^t)
    float sqrt_bc2;                   // sqrt(1 - beta2^t)
};

// Workgroups of the optimizer kernel: one per 32x32 weight tile, one per four biases.
__host__ __device__ inline int opt_tiles(int out, int kp) { return (out / 32) * (kp / 32); }
__host__ __device__ inline int opt_bias_groups(int out) { return out / 4; }
hipError_t optimizer(const OptArgs& a, int n_groups, hipStream_t stream);

// Network-scored differential evolution (de.hip).  x [pop][dim] the population,
// f [pop] its objectives, X [pop_pad][n_in] the network input (trials in the first
// dim columns, the fixed inputs after them; the propose kernel also writes it packed
// into X0, as pack_input would), Y [pop_pad][n_out] the network output,
// keys [pop] each member's (objective bits << 32 | member) after the last scoring.
struct DeArgs {
    float* x;                         // members read by this generation [pop][dim]
    float* x_next;                    // members after its selection
    float* f;                         // objectives [pop] (member i's only by its own selection)
    float* X;
    __bf16* X0;                       // the network's packed input [pop_pad][kp0] (BF16)
    int kp0;                          // its row length (n_in padded for the K tiling)
    const float* Y;
    const float* target;              // [n_out]
    const unsigned long long* keys;   // argmin keys of x [pop]
    unsigned long long* keys_next;    // of x_next
    int pop, pop_pad, dim, n_in, n_out;
    float F, CR;
    float log_span, out_scale, inv_peak2;
    uint32_t key0, key1;
    float fixed[8];
};
hipError_t de_init(const DeArgs& a, hipStream_t stream);
hipError_t de_propose(const DeArgs& a, uint32_t gen, hipStream_t stream);
hipError_t de_score(const DeArgs& a, hipStream_t stream);

// The network as the fused generation kernel reads it: per layer the BF16 weights
// in fragment order (de_pack of the [out][kp] weights) and the FP32 bias; lda =
// the LDS row stride of its activation buffers (the widest kp or hidden width,
// plus 8).
struct DeLayer {
    const __bf16* w;
    const float* b;
    int out, kp;
};
struct DeNet {
    DeLayer L[kMaxLayers];
    int n_layers, lda;
};
hipError_t de_pack(const __bf16* W, int out, int kp, __bf16* Wf, hipStream_t stream);
int de_fused_lds(const DeNet& net, int n_out);
hipError_t de_fused_prepare(int lds_bytes);
hipError_t de_generation_fused(const DeArgs& a, const DeNet& net, uint32_t gen, int lds_bytes,
                               hipStream_t stream);

// The inference forward in one kernel (mlp_fused.hip): 128 rows per workgroup
// through every layer, activations in LDS, weights in fragment order (de_pack).
// Hidden layers 512 wide (the first with its inputs padded to 64), the output
// layer 256 or 512 wide.
bool infer_fused_ok(const DeNet& net);
int infer_fused_lds(const DeNet& net);
hipError_t infer_fused_prepare(int lds_bytes);
hipError_t infer_fused(const float* x, int B, int n_in, const DeNet& net, float* y, int n_out,
                       int lds_bytes, hipStream_t stream);

}  // namespace mlp
}  // namespace mcl
