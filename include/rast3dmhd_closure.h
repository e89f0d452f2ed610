/*
 * rast3dmhd_closure.h -- the learned subgrid closure's network.
 *
 * No counterpart in the reference (which has no closure).  A network of
 * three 3x3x3 convolutions, cin -> width -> width -> cout, GELU after the
 * first two, maps filtered coarse-grid fields to subgrid flux divergences.
 * Weights and activations are BF16 on MFMA with FP32 accumulation; the
 * master parameters and the Adam moments are FP32 on the device.
 *
 * Volumes are nz x ny x nx points, channel-last float arrays
 * x[z][y][x][c] on the host, periodic in x and y, zero padding in z.
 * Parameters are one FP32 vector in the order W1, b1, W2, b2, W3, b3 with
 * W[co][tap][ci], tap = 9 (dz+1) + 3 (dy+1) + (dx+1).
 * Return codes are those of rast3dmhd.h.
 */
#ifndef RAST3DMHD_CLOSURE_H
#define RAST3DMHD_CLOSURE_H

#include "rast3dmhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd_net mhd_net_t;

/* Built for cin = 8, width = 32, cout <= 8 (MHD_ERR_SCOPE otherwise).
 * Parameters start at zero: set them with mhd_net_set_params. */
int mhd_net_create(int cin, int width, int cout, mhd_net_t **out);
long long mhd_net_num_params(const mhd_net_t *n);
/* Also resets the Adam moments. */
int mhd_net_set_params(mhd_net_t *n, const float *params);
int mhd_net_get_params(mhd_net_t *n, float *params);

/* y = net(x); x has cin channels, y cout channels. */
int mhd_net_forward(mhd_net_t *n, int nx, int ny, int nz, const float *x, float *y);

/* Gradients of sum(dy * net(x)) with respect to the parameters (dparams,
 * nparams floats) and, if dx is not NULL, the input.  dy is rounded to
 * BF16 where the kernels read it. */
int mhd_net_gradient(mhd_net_t *n, int nx, int ny, int nz, const float *x, const float *dy, float *dparams,
                     float *dx);

/* Training set resident on the device: input x (cin channels) and target t
 * (cout channels) per sample. */
int mhd_net_add_sample(mhd_net_t *n, int nx, int ny, int nz, const float *x, const float *t);
int mhd_net_num_samples(const mhd_net_t *n);
int mhd_net_clear_samples(mhd_net_t *n);

/* iters Adam steps (beta 0.9, 0.999) on the mean squared error, cycling
 * through the samples; the losses before the first and the last step. */
int mhd_net_train(mhd_net_t *n, int iters, float lr, double *loss_first, double *loss_last);

/* Mean squared error of the current parameters on sample s. */
int mhd_net_loss(mhd_net_t *n, int s, double *loss);

/* Mean milliseconds of one forward pass over an nx x ny x nz volume
 * (HIP events; reps passes after one warm-up). */
int mhd_net_time_forward(mhd_net_t *n, int nx, int ny, int nz, int reps, double *ms);

/* Closure-corrected step of a MHD_MIXED box (mhd_t of rast3dmhd.h): every
 * following step evaluates the network on the state at its start and adds
 * its output to the right-hand sides of RU, RV, RW, TT, BX, BY, BZ in all
 * three stages.  Inputs (rho, u, v, w, s = ln T / (gamma - 1) - ln rho, BX,
 * BY, BZ) enter as (value - in_profile[c][k]) / in_std[c]; the forcing is
 * output * out_std[c] + out_profile[c][k], where the profiles hold one value
 * per interior plane k = 0..NPZ-1 (in_profile [8][NPZ], out_profile [7][NPZ]).
 * Fallback: with net NULL, or val_error (the network's error on held-out
 * data) above threshold, the box steps plainly; a step whose network output
 * is not finite is also plain.  The network must live on the box's device
 * and stay alive while attached. */
int mhd_set_closure(mhd_t *h, mhd_net_t *net, const double *in_profile, const double *in_std,
                    const double *out_profile, const double *out_std, double val_error, double threshold);
/* 1 if a network is attached (mhd_set_closure accepted it), else 0. */
int mhd_closure_active(const mhd_t *h);

const char *mhd_net_last_error(const mhd_net_t *n);
void mhd_net_destroy(mhd_net_t *n);

#ifdef __cplusplus
}
#endif
#endif
