/* C ABI of the learned subgrid-stress (SGS) closure: an LES on a coarse grid
 * runs the same pseudo-spectral RK2 step as the DNS (include/psdns_gpu.h,
 * per-routine entry points) with a predicted subgrid stress tau_ij added to
 * the physical-space products, and the closure is trained on box-filtered
 * fields of the GPU DNS itself.
 *
 * No counterpart in the reference: PSDNS is a DNS code and has no LES
 * closure.  These entry points serve the learned part that the build's
 * brief adds beside the DNS hot path.
 *
 * Model (parameters theta, FP32, psdns_sgs_param_count() of them):
 *   features per LES grid point, from the velocity gradient A = du_i/dx_j
 *   of the masked spectral state (i k_j, inverse transforms of kxtran):
 *     S = sym(A) - tr(A)/3 I, W = skew(A), q = |S|^2 + |W|^2,
 *     lambda = (tr S^2 / q, tr W^2 / q, tr S^3 / q^1.5, tr W^2 S / q^1.5,
 *               tr W^2 S^2 / q^2)
 *     T = delta^2 (|S| S, S W - W S, S^2 - tr S^2/3 I, W^2 - tr W^2/3 I),
 *     delta = 2 pi / n (the LES grid spacing), |S| = sqrt(2 S:S);
 *   network: h0 = tanh(conv3x3x3(lambda) 5 -> 16 channels, periodic),
 *     h1..h3 = tanh(dense 64) on [lambda, h0], g = dense 4 of h3;
 *   stress: tau_hat = sum_n g_n T_n (symmetric, traceless: the trace is a
 *     pressure and drops out of the projected step).
 *   theta = W0[16][135] b0[16] W1[64][21] b1[64] W2[64][64] b2[64]
 *           W3[64][64] b3[64] W4[4][64] b4[4]  (conv input k = ch*27 +
 *           ((dy+1)*3 + dz+1)*3 + dx+1).
 * Target: tau_ij = filter(u_i u_j) - filter(u_i) filter(u_j), deviatoric,
 *   the filter a top-hat of width r fine cells (trapezoidal end weights)
 *   at every r-th fine point.
 * Loss: sum over points of sum_ij (tau_hat - tau)_ij^2 / (points sigma2).
 * Fields are physical [c][y][z][x], n^3 floats per component; symmetric
 * tensors have 6 components (11, 12, 13, 22, 23, 33).
 *
 * Errors, threading and synchronisation as psdns_gpu.h: 0 or a negative
 * PSDNS_E* code, psdns_last_error() holds the message; one host thread per
 * object; every call has finished its device work when it returns.
 */
#ifndef PSDNS_SGS_H
#define PSDNS_SGS_H

#include "psdns_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psdns_sgs psdns_sgs;

/* fields of psdns_sgs_get_field */
enum {
  PSDNS_SGS_LAMBDA = 0,    /* 5 invariants */
  PSDNS_SGS_BASIS = 1,     /* 4 x 6 basis tensors T[n][c] */
  PSDNS_SGS_TARGET = 2,    /* 6: the filtered DNS stress (psdns_sgs_filter) */
  PSDNS_SGS_PREDICTED = 3, /* 6: the network's stress (psdns_sgs_predict) */
  PSDNS_SGS_GRADIENT = 4   /* 9: A[3 i + j] = du_i/dx_j */
};

int psdns_sgs_param_count(void);
/* a closure for the LES grid n^3 with room for max_snapshots training
 * snapshots (35 n^3 floats each) */
int psdns_sgs_create(int n, int max_snapshots, psdns_sgs **out);
int psdns_sgs_destroy(psdns_sgs *m);
/* parameters (count = psdns_sgs_param_count()); set resets the Adam state */
int psdns_sgs_set_params(psdns_sgs *m, const float *theta, int count);
int psdns_sgs_get_params(psdns_sgs *m, float *theta, int count);

/* Box filter of the fine context's physical velocity (ux slots 0..2, e.g.
 * after psdns_kxtran(fine, 3)) at ratio r = n_fine / n_coarse (even):
 * the filtered velocity into the coarse context's ux slots 0..2 (then
 * psdns_xktran(coarse, 3) gives its spectral state) and the deviatoric
 * stress as the closure's target.  Both contexts in one precision. */
int psdns_sgs_filter(psdns_sgs *m, psdns_ctx *fine, psdns_ctx *coarse);
/* Features of the context's cylinder u slots 0..2 as they are: unshifted
 * after psdns_xktran, or on the stage's shifted grid after psdns_phase1 */
int psdns_sgs_features(psdns_sgs *m, psdns_ctx *c);
/* Features of the context's state u on stage kstep's shifted grid (1|2:
 * the phase shift of psdns_set_shifts applied on the fly, as the fused
 * step's phase1): for an LES stage through the fused entry points */
int psdns_sgs_features_stage(psdns_sgs *m, psdns_ctx *c, int kstep);
/* LES stage through the fused entry points: the predicted stress is added
 * to the products by the next psdns_stage_nonlinear(les, kstep) (in its x
 * pass; used once, dropped by any call that overwrites the work arrays or
 * the shifts).  The closure must stay alive until then.  A context with
 * passive scalars is refused (PSDNS_EINVAL). */
int psdns_sgs_attach_stress(psdns_sgs *m, psdns_ctx *les, int kstep);
/* host data as the current features (5 + 24 fields) and target (6; nullable) */
int psdns_sgs_put_fields(psdns_sgs *m, const float *lambda, const float *basis, const float *tau);
int psdns_sgs_get_field(psdns_sgs *m, int which, float *out);

/* training store: the current features and target as one snapshot */
int psdns_sgs_add_snapshot(psdns_sgs *m);
int psdns_sgs_snapshots(psdns_sgs *m, int *count);
int psdns_sgs_clear_snapshots(psdns_sgs *m);
/* sigma2 = mean over stored points of sum_ij tau_ij^2 (FP64) */
int psdns_sgs_target_scale(psdns_sgs *m, double *sigma2);

/* Minibatch of `batch` stored points (the index map: point g =
 * splitmix64(base + b) mod (snapshots n^3), base = splitmix64(seed *
 * 0xD1B54A32D192ED03 + step)): loss and dloss/dtheta in FP32 without an
 * update */
int psdns_sgs_gradient(psdns_sgs *m, int batch, unsigned long long seed, long long step, double sigma2,
                       float *grad, double *loss);
/* nsteps Adam steps (lr, beta 0.9 / 0.999, eps 1e-8), step t using the
 * minibatch of (seed, t); losses[nsteps] (nullable) the minibatch losses */
int psdns_sgs_train(psdns_sgs *m, int nsteps, int batch, double lr, unsigned long long seed, double sigma2,
                    double *losses);
/* predicted stress of the current features: the convolution in FP32, then
 * the dense layers in FP32 (bf16 = 0) or with bf16 operands on the matrix
 * cores and FP32 accumulation (bf16 = 1) */
int psdns_sgs_predict(psdns_sgs *m, int bf16);
/* loss of the prediction against the current target over all points */
int psdns_sgs_eval(psdns_sgs *m, double sigma2, double *loss);
/* LES stage: the predicted stress added to the products in the context's
 * ux slots 0..4 (after psdns_realspace_vel, before psdns_transform):
 * P1 += tau11 - tau22, P2 += tau12, P3 += tau33 - tau22, P4 += tau23,
 * P5 += tau13 */
int psdns_sgs_add_stress(psdns_sgs *m, psdns_ctx *les);

#ifdef __cplusplus
}
#endif

#endif
