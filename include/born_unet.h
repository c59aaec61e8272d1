/*
 * born_unet.h -- learned image-domain preconditioner of the MI355X Born
 * RTM library.  It has no counterpart in the reference, whose inversion
 * layer only offers an operator-chain hook (prec_solver,
 * Born/lin_solver.cpp:21-37) and plain cgstep.
 *
 * Network: a 3-level 3-D U-Net on P^3 patches (P a multiple of 64), 3x3x3
 * convolutions with 16 / 32 / 64 channels, ReLU, 2x2x2 average pooling,
 * nearest upsampling with skip concatenation, a 1x1x1 head and a residual
 * connection (out = x + head(...), so an untrained network is close to the
 * identity).  Convolutions are implicit GEMMs in BF16 on MFMA with FP32
 * accumulation; master weights, gradients and Adam state are FP32.
 *
 * As a preconditioner the network maps a model-space gradient g to
 * z = s * net(g / s), s = RMS(g), applied on P^3 patches that overlap by
 * `overlap` cells on each side (the central (P - 2 overlap)^3 of each patch
 * is kept).  Training pairs are (RTM image, LSRTM image) made by this
 * library's own operator and CG.
 */
#ifndef BORN_UNET_H
#define BORN_UNET_H

#include "born_op.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct brtm_unet brtm_unet;

/* batch: patches per network pass; patch: P (multiple of 64). */
int brtm_unet_create(brtm_ctx *ctx, int batch, int patch, unsigned long long seed,
                     brtm_unet **out);
int brtm_unet_destroy(brtm_unet *net);
long brtm_unet_num_params(const brtm_unet *net);
int brtm_unet_get_params(brtm_unet *net, float *host);
int brtm_unet_set_params(brtm_unet *net, const float *host);

/* y = net(x) for nb <= batch patches; x, y are device FP32 [nb][P^3]. */
int brtm_unet_forward(brtm_unet *net, const float *x, float *y, int nb);
/* loss = mean((net(x) - t)^2); gradients of the loss w.r.t. the FP32
 * parameters into grad (device or host pointer by grad_on_host). */
int brtm_unet_loss_grad(brtm_unet *net, const float *x, const float *t, int nb, double *loss,
                        float *grad, int grad_on_host);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the MSE loss. */
int brtm_unet_train_step(brtm_unet *net, const float *x, const float *t, int nb, float lr,
                         double *loss);

/* Copy the P^3 patch of a volume [m3][m2][m1] (device FP32) with origin
 * (o1,o2,o3) (may lie partly outside: zeros) into dst (device FP32 P^3),
 * scaled by scale. */
int brtm_unet_extract(brtm_unet *net, const float *vol, int m1, int m2, int m3, int o1, int o2,
                      int o3, float scale, float *dst);

/* z = s net(g / s) on overlapping patches of the model grid. */
int brtm_unet_apply(brtm_unet *net, const brtm_vec *g, brtm_vec *z, int m1, int m2, int m3,
                    int overlap);
/* Model grid and overlap used by brtm_unet_precond. */
int brtm_unet_set_grid(brtm_unet *net, int m1, int m2, int m3, int overlap);
/* Split weight W (a positive model-grid diagonal, e.g. an inverse square
 * root of the source illumination; NULL clears it): brtm_unet_precond then
 * returns z = W P(W g), so that P works on amplitude-balanced gradients and
 * W P W is the illumination-diagonal preconditioner when P is the identity.
 * A network output that is not a descent direction (z.g <= 0 or not finite)
 * is replaced by W^2 g, and so is one that stopped following its input (see
 * brtm_unet_set_guard).  The vector is not copied: keep it alive while it is
 * set.  Setting it (or the grid) starts a new solve for the guard. */
int brtm_unet_set_split_weight(brtm_unet *net, const brtm_vec *w);
/* Thresholds of the split preconditioner's guard (defaults 0.8 and 0.05).
 * min_cos: a z with cos(z, g) < min_cos brings nothing the gradient does
 * not, and becomes W^2 g.  repeat_cos: a network output (mode 0) or
 * correction (mode 1, the default) whose cosine with the previous one
 * exceeds it has settled on a fixed direction, which stalls flexible CG;
 * brtm_unet_set_guard_mode says what each mode then does.
 * brtm_unet_guard_count returns how many calls the guard acted on since the
 * split weight was set. */
int brtm_unet_set_guard(brtm_unet *net, double repeat_cos, double min_cos);
int brtm_unet_guard_count(const brtm_unet *net);
/* Guard mode 1 (the default): z = W^2 g + c, c = the network's correction
 * (its split output minus W^2 g); when c repeats the previous correction
 * (cos above the repeat threshold) only its component orthogonal to that one
 * is kept, and a z that is no descent direction or nearly orthogonal to g
 * falls back to W^2 g.  brtm_unet_guard_count counts both.  Mode 0 replaces
 * a repeating network output by W^2 g outright.  No counterpart in the
 * reference. */
int brtm_unet_set_guard_mode(brtm_unet *net, int mode);
/* The preconditioner uses the odd part of the network,
 * P(x) = (net(x) - net(-x)) / 2, on = 1 (the default for batch >= 2): the
 * ideal preconditioner is linear, hence odd, and the even part -- a
 * response of the biases that does not follow the input -- otherwise turns
 * into a fixed direction that stalls the solver after a few iterations.
 * Costs a second network evaluation per patch. */
int brtm_unet_set_odd(brtm_unet *net, int on);
/* brtm_precond_fn for brtm_cg_solve: user = brtm_unet*.  The solver's
 * gradient g = L^T(L m - d) is minus the RTM image of the data residual, so
 * the network is applied in its training orientation:
 * z = -s net(-g / s), s = rms(g). */
int brtm_unet_precond(void *user, const brtm_vec *g, brtm_vec *z);

#ifdef __cplusplus
}
#endif
#endif /* BORN_UNET_H */
