/*
 * mcl_surrogate.h — C ABI of the learned glow-curve surrogate in libmcl.so.
 *
 * A multilayer perceptron (n_in -> n_hidden x n_hidden_layers, SiLU -> n_out)
 * whose matrix products run on gfx950 BF16 matrix cores with FP32 accumulation,
 * trained by the library's own forward, backward and AdamW kernels with FP32
 * master weights.  It maps trap parameters to a binned glow or decay curve and
 * stands in for the Monte Carlo inside parameter fits.
 *
 * The reference has no surrogate: its fit (src/class/optimizer.py:32-111) calls
 * the simulation for every candidate through run_one_sim / objective
 * (optimizer.py:66-84).  These entry points replace that per-candidate
 * simulation by a batched network evaluation; plain Monte Carlo
 * (mcl_run_fixed_dt, include/mcl.h) stays the fallback and the judge.
 *
 * Conventions (as in include/mcl.h): status MCL_OK or a negative MCL_E* code,
 * message in mcl_last_error(ctx); the network runs on the context's stream.
 * Batch data (x, t, idx, y, grad, loss) are DEVICE pointers; parameter
 * import/export uses HOST pointers.
 *
 * Parameter layout (the "flat" vector, FP32): for each linear layer l in order,
 * W_l as [out_l][in_l] row-major, then b_l [out_l].  Layer 0 maps n_in -> n_hidden,
 * layers 1 .. n_hidden_layers-1 map n_hidden -> n_hidden, the last maps
 * n_hidden -> n_out.  y = W x + b, SiLU after every layer but the last.
 */
#ifndef MCL_SURROGATE_H
#define MCL_SURROGATE_H

#include <stdint.h>

#include "mcl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcl_mlp_desc {
    int32_t n_in;             /* inputs (at most 64)                              */
    int32_t n_hidden;         /* hidden width (multiple of 128)                   */
    int32_t n_hidden_layers;  /* SiLU layers (at least 1)                         */
    int32_t n_out;            /* outputs (multiple of 128)                        */
    int32_t max_batch;        /* largest batch of any call (multiple of 64)       */
} mcl_mlp_desc;

/* AdamW hyper-parameters of one optimizer step (decoupled weight decay,
 * bias-corrected moments; step counts from 1). */
typedef struct mcl_adamw {
    double lr, beta1, beta2, eps, weight_decay;
    int64_t step;
} mcl_adamw;

int     mcl_mlp_create(void* ctx, const mcl_mlp_desc* desc, void** mlp_out);
void    mcl_mlp_destroy(void* mlp);
int64_t mcl_mlp_param_count(const void* mlp);
/* FP32 master parameters in the flat layout (host pointers). Setting them
 * also refreshes the BF16 copies and clears the AdamW moments. */
int     mcl_mlp_get_params(void* mlp, float* params);
int     mcl_mlp_set_params(void* mlp, const float* params);

/* Inference: y[b][n_out] = net(x[b][n_in]), FP32 in and out, batch a multiple
 * of 64 and at most max_batch. */
int mcl_mlp_forward(void* mlp, const float* x, int32_t batch, float* y);

/*
 * Loss and gradient of L = mean over batch rows r and outputs j of
 * (net(x[i_r])_j - t[i_r][j])^2, with i_r = idx[r] (idx NULL: i_r = r).
 * grad: flat layout [param_count] FP32, or NULL; loss: one FP32, or NULL.
 * Batch a multiple of 64 and at most max_batch.
 */
int mcl_mlp_grad(void* mlp, const float* x, const float* t, const int32_t* idx,
                 int32_t batch, float* grad, float* loss);

/* One training step: the gradient above followed by an AdamW update of the
 * master weights and their BF16 copies, without a host synchronisation. */
int mcl_mlp_train_step(void* mlp, const float* x, const float* t, const int32_t* idx,
                       int32_t batch, const mcl_adamw* hp, float* loss);

/* Device time of the last forward / grad / train call (HIP events). */
int mcl_mlp_last_ms(void* mlp, double* ms);

/*
 * Network-scored differential evolution: the search of the surrogate-assisted
 * fit.  Replaces the per-candidate objective calls of the reference's scipy DE
 * (optimizer.py:98-111, objective optimizer.py:66-84) while the network scores
 * the candidates: the population lives in device memory and a generation is one
 * propose kernel (DE/rand-to-best/1/bin: mutant = x_a + F (x_best - x_a) +
 * F (x_b - x_c) over three distinct other members, binomial crossover with one
 * forced coordinate, coordinates outside [-1, 1] redrawn uniformly;
 * Philox4x32-10 draws keyed by seed), the network forward, and one
 * score-and-accept kernel: objective = mean over the n_out outputs of
 * (expm1(y log_span) out_scale - target)^2 * inv_peak2; a trial replaces its
 * member when not worse; the argmin goes to the lowest member on ties.
 * The searched coordinates are the first dim network inputs, in [-1, 1]; the
 * other n_in - dim inputs are the fixed values.  A search runs the network on
 * the network's own activation buffers: calls of the same network (another
 * search, forward, grad, train) must not overlap it on another stream.
 */
typedef struct mcl_de_params {
    int32_t pop;              /* members (at least 4; rounded up to 64, at most max_batch) */
    int32_t dim;              /* searched inputs (1 .. 8, at most n_in)                    */
    float F, CR;              /* mutation scale [0, 2], crossover probability [0, 1]      */
    uint64_t seed;            /* Philox key                                                */
    float fixed[8];           /* inputs dim .. n_in-1 (n_in - dim <= 8)                    */
    float log_span, out_scale, inv_peak2;
    int32_t flags;            /* MCL_DE_LAYERED: a generation as separate kernels          */
} mcl_de_params;

/* mcl_de_params.flags: run each generation as a propose kernel, the network's
 * per-layer GEMM kernels and a score kernel instead of one fused kernel (the
 * results are identical; the fused kernel needs layers at most 512 wide). */
#define MCL_DE_LAYERED 1

/* Creates the search and scores its initial population (uniform in [-1, 1]);
 * target: DEVICE [n_out], read by every generation.  The fused generation kernel
 * reads a copy of the network's weights made here: training the network during
 * the search does not change what the search sees. */
int  mcl_de_create(void* mlp, const mcl_de_params* params, const float* target, void** de_out);
void mcl_de_destroy(void* de);
/* Runs gens generations on the context's stream, without a host synchronisation. */
int  mcl_de_run(void* de, int32_t gens);
/* The best member (x HOST [dim]) and its objective (f HOST); synchronises. */
int  mcl_de_best(void* de, float* x, float* f);
/* The population (x DEVICE [pop][dim]) and objectives (f DEVICE [pop]), copied on the stream. */
int  mcl_de_population(void* de, float* x, float* f);

#ifdef __cplusplus
}
#endif
#endif /* MCL_SURROGATE_H */
