/* libdgk surrogate: a Fourier-feature MLP mapping (source xyz, receiver xyz) to a receiver
 * pressure time series, trained and evaluated with hand-written gfx950 kernels (BF16 MFMA GEMMs
 * with FP32 accumulation, fused bias + SiLU epilogues, relative-L2 loss, Adam).
 *
 * No counterpart in the reference (povanberg/DGFEM-Acoustic has no learned component); it
 * stands beside the DG solve, which generates its training data (dgk_batch_* with receivers) and
 * answers the queries the surrogate declines.
 *
 * Network: z = [sin(2 pi B x), cos(2 pi B x)] (B: n_freq x in_dim, fixed), depth hidden layers of
 * `width` units with SiLU, a linear output layer of out_dim values. Parameters are one flat FP32
 * vector [W_1 b_1 ... W_L b_L] with W_l row-major [out][in]; L = depth + 1. */
#ifndef DGK_SURROGATE_H
#define DGK_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { DGS_OK = 0, DGS_ERR_INVALID = 1, DGS_ERR_DEVICE = 4 } dgs_status;

typedef struct {
  int in_dim;        /* input coordinates (6: source xyz, receiver xyz, normalised to [0,1]) */
  int n_freq;        /* Fourier frequencies (multiple of 32); 2*n_freq features */
  int width;         /* hidden width (multiple of 64) */
  int depth;         /* hidden layers */
  int out_dim;       /* output samples (multiple of 64) */
  int max_batch;     /* rows per device pass (multiple of 64) */
  float freq_sigma;  /* standard deviation of the Fourier matrix entries */
  uint64_t seed;     /* initialisation seed (weights He-normal, biases 0, Fourier matrix) */
} dgs_config;

typedef struct dgs_net dgs_net;

dgs_status dgs_create(const dgs_config* cfg, dgs_net** out);
dgs_status dgs_destroy(dgs_net* net);
const char* dgs_last_error(const dgs_net* net);
int64_t dgs_num_params(const dgs_net* net);
/* copy parameters (num_params floats) and/or the Fourier matrix (n_freq*in_dim); either may be NULL.
 * Setting parameters resets the optimizer state. */
dgs_status dgs_get_params(dgs_net* net, float* params, float* fourier);
dgs_status dgs_set_params(dgs_net* net, const float* params, const float* fourier);
/* y[rows][out_dim] = net(x[rows][in_dim]) (host buffers, any row count) */
dgs_status dgs_forward(dgs_net* net, int rows, const float* x, float* y);
/* y[rows][out_dim] = net(x[rows][in_dim]) with x and y in device memory of the network's GPU.
 * No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): it calls no allocating routine; its activations live
 * in a fixed 16 MiB device array per GPU (rows are processed in passes of the largest multiple
 * of 64 whose buffers fit; a network whose 64-row pass does not fit is rejected with
 * DGS_ERR_INVALID). Ordered on the network's stream and synchronous like dgs_forward. Calls
 * that overlap on one GPU share that array and race. */
dgs_status dgs_forward_noalloc(dgs_net* net, int rows, const float* x, float* y);
/* average device time of `reps` forward passes over `rows` rows already on the device */
dgs_status dgs_forward_timed(dgs_net* net, int rows, int reps, double* ms_per_rep);
/* one forward + backward without update: loss = sum_r w[r] |y_r - t_r|^2 and its gradient
 * (num_params floats, may be NULL); rows a multiple of 64 up to max_batch */
dgs_status dgs_gradients(dgs_net* net, int rows, const float* x, const float* t, const float* w, float* grad,
                         double* loss);
/* keep a training set resident on the device: x[rows][in_dim], t[rows][out_dim], w[rows] */
dgs_status dgs_set_data(dgs_net* net, int64_t rows, const float* x, const float* t, const float* w);
/* `steps` Adam steps on minibatches of `batch` rows (seeded epoch permutations of the resident
 * set); minibatch loss = mean over rows of w_r |y_r - t_r|^2; losses[steps] may be NULL */
dgs_status dgs_train(dgs_net* net, int steps, int batch, float lr, uint64_t seed, float* losses);

#ifdef __cplusplus
}
#endif

#endif
