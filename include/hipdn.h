/* hipdn: kernels of the 3D convolutional Dij-column denoiser on one MI355X (gfx950).
 *
 * The reference has no learned component: these entry points have no counterpart in
 * it. They serve the denoised Dij builder (pyradmc_amd/denoise.py), which predicts
 * full-statistics beamlet columns from low-history ones and falls back to plain Monte
 * Carlo for any column that fails its consistency check.
 *
 * Conventions. Every pointer argument is device memory owned by the caller; `stream`
 * is a hipStream_t (NULL = the null stream). Activations are bfloat16, position-major
 * ("nd": [N][D][H][W][C], C a multiple of 16) with an optional channel-major copy
 * ("cm": [C][N*D*H*W]). Convolutions are 3x3x3, stride 1, zero padding. Weights: FP32
 * master [27][Cout][Cin]; BF16 forward copy [27][Cout][Cin]; BF16 data-gradient copy
 * [27][Cin][Cout] with flipped taps. Calls return HDN_OK or an error code and are
 * asynchronous on `stream`.
 */
#ifndef HIPDN_H
#define HIPDN_H

#ifdef __cplusplus
extern "C" {
#endif

#define HDN_OK 0
#define HDN_EINVAL -1
#define HDN_EHIP -2
#define HDN_ESIZE -3 /* the call needs more than the library's fixed device array holds */

/* Floats of the fixed per-GPU device array behind hdn_conv3d_wgrad_noalloc (64 MiB). */
#define HDN_SCRATCH_FLOATS (16LL * 1024 * 1024)

/* y = act(bias + conv(x, w)) with act(v) = v >= 0 ? v : slope v (slope 1: identity);
 * y_cm may be NULL; bias may be NULL. N*D*H*W % 64 == 0, cin % 16 == 0,
 * cout % 16 == 0. With w = the data-gradient copy, x = dz and bias = NULL, slope = 1
 * this is the data gradient of a convolution. */
int hdn_conv3d_fwd(const void* x, const void* w, const float* bias, void* y_nd, void* y_cm, int n,
                   int d, int h, int w_, int cin, int cout, float slope, void* stream);
/* The same convolution reading its input through the preceding element-wise step, for
 * inference (the values equal those of the separate calls, so the outputs are
 * identical): src 0 reads x; src 1 (HDN_SRC_UP) reads hdn_up_fwd(x), x having half
 * the output dims d, h, w (all even); src 2 (HDN_SRC_ADD) reads hdn_add(x, x2); src 3
 * (HDN_SRC_LOW8) reads x whose channels beyond the first 8 are zero (the padded network
 * input), skipping them in the reduction. */
#define HDN_SRC_PLAIN 0
#define HDN_SRC_UP 1
#define HDN_SRC_ADD 2
#define HDN_SRC_LOW8 3
int hdn_conv3d_fwd_src(const void* x, const void* x2, int src, const void* w, const float* bias, void* y_nd,
                       void* y_cm, int n, int d, int h, int w_, int cin, int cout, float slope, void* stream);

/* Weight-gradient partials part[splits][27][cout][cin] (FP32) from the channel-major
 * layer input and output gradient; W % 8 == 0. Sum them with hdn_reduce_splits. */
int hdn_conv3d_wgrad(const void* x_cm, const void* dz_cm, float* part, int splits, int n, int d,
                     int h, int w_, int cin, int cout, void* stream);

/* No counterpart in the reference (which has no learned component): the weight gradient
 * for callers that cannot allocate device memory - code under HIP graph capture, training
 * loops and harnesses that forbid hipMalloc. grad[27][cout][cin] (FP32) receives what
 * hdn_conv3d_wgrad followed by hdn_reduce_splits would give with the same splits, bit
 * for bit; the partials go to a fixed device array the library holds on each GPU until
 * the process ends, so nothing is allocated. Limit: splits * 27 * cout * cin must not
 * exceed HDN_SCRATCH_FLOATS, else HDN_ESIZE and nothing runs. `stream` must belong to
 * the current device (else HDN_EINVAL). Every call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU (other streams, other host threads) race. The
 * first call on a device looks up the array's address there. */
int hdn_conv3d_wgrad_noalloc(const void* x_cm, const void* dz_cm, float* grad, int splits, int n,
                             int d, int h, int w_, int cin, int cout, void* stream);

/* out[i] = sum over s of part[s][i], in split order (deterministic). */
int hdn_reduce_splits(const float* part, float* out, long long len, int splits, void* stream);

/* Per-channel sums of a channel-major tensor over its P positions: part[splits][C]. */
int hdn_channel_sum(const void* cm, float* part, long long P, int C, int splits, void* stream);

/* In the element-wise calls below, the channel-major copy (y_cm, dz_cm) may be NULL:
 * it feeds only the weight gradients, so inference skips it. */

/* dz = dy * act'(y) for the leaky activation (from the stored output y). */
int hdn_act_bwd(const void* y, const void* dy, void* dz_nd, void* dz_cm, long long P, int C,
                float slope, void* stream);

/* y = a + b. */
int hdn_add(const void* a, const void* b, void* y_nd, void* y_cm, long long P, int C, void* stream);

/* 2x2x2 average pooling to the coarse dims (n, d, h, w) and its gradient (fine dims). */
int hdn_pool_fwd(const void* x, void* y_nd, void* y_cm, int n, int d, int h, int w_, int C,
                 void* stream);
int hdn_pool_bwd(const void* dy, void* dx, int n, int d2, int h2, int w2, int C, void* stream);

/* Nearest 2x upsampling to the fine dims and its gradient (coarse dims). */
int hdn_up_fwd(const void* x, void* y_nd, void* y_cm, int n, int d2, int h2, int w2, int C,
               void* stream);
int hdn_up_bwd(const void* dy, void* dx, int n, int d, int h, int w_, int C, void* stream);

/* Residual head and weighted squared error: pred = inp[:, 0] + out[:, 0];
 * partial[block] = sum weight (pred - target)^2 (blocks of 256 positions);
 * dout = 2 weight (pred - target) grad_scale in channel 0 (others 0). target, weight,
 * dout_nd, dout_cm, pred and partial may each be NULL (prediction only). */
int hdn_loss(const void* out, const void* inp, int c_out, int c_in, const float* target,
             const float* weight, void* dout_nd, void* dout_cm, float* pred, float* partial,
             long long P, float grad_scale, void* stream);

/* BF16 forward and data-gradient copies of a FP32 master weight. */
int hdn_prepare_weights(const float* master, void* wf, void* wd, int cin, int cout, void* stream);

/* One Adam step (bias-corrected; step counts from 1) with decoupled L2 term
 * weight_decay * param added to the gradient. */
int hdn_adam(float* param, const float* grad, float* m, float* v, long long len, float lr,
             float beta1, float beta2, float eps, int step, float weight_decay, void* stream);

#ifdef __cplusplus
}
#endif
#endif
