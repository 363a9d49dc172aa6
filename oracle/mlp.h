/*
 * Learned Metropolis-Hastings proposal: CPU restatement (test infrastructure
 * only -- the oracle of the device MLP kernels, never part of the product).
 *
 * No counterpart in the reference (its proposals are MetropolisBox and
 * MetropolisDiffuse, src/metropolis/src/metrop.rs:21-212); the proposal plugs
 * into the same propose/accept contract (metropolis/src/traits.rs:10-38).
 *
 * Network: 13 -> 64 -> 64 -> 64 -> 4, tanh on the hidden layers.
 *   inputs of electron i at r (other electrons at their positions):
 *     f[0]      distance to the nearest nucleus
 *     f[1..9]   x_j - r for the three nearest other electrons, nearest first
 *               (ties: lower index first; missing partners give zeros)
 *     f[10..12] r
 *   outputs: mu = out[0..2], ls = ln(sigma) = LS0 + 3 tanh(out[3] / 3)
 * Proposal: x' = (r + mu) + sigma z, z ~ N(0, I3) (Box-Muller draws as the
 * diffusion move).  log q(x'|x) = -3 ls - |z|^2/2, log q(x|x') = -3 ls' -
 * |(r - x') - mu'|^2 / (2 sigma'^2) with (mu', ls') the network at x'; the
 * move is accepted iff min(R^2 exp(log q(x|x') - log q(x'|x)), 1) > u, which
 * is the exact MH ratio for this q.
 * Parameters (FP32, 9476): W1[13][64] b1[64] W2[64][64] b2[64] W3[64][64]
 * b3[64] W4[64][4] b4[4], z = b + sum_in x_in W[in][out].
 * Training objective per sample (objective 0): O = A |x' - r|^2 (expected
 * squared jump distance times acceptance); its gradient treats the inputs of
 * the reverse evaluation f(x') as constants (their dependence on x' is not
 * propagated).  Objective 1 divides O by d^2 + 0.1^2, d the larger of the
 * float-rounded nearest-nucleus distances f[0] at r and at x' (a constant of
 * the gradient); objective 2 is log(1e-4 + O), whose gradient is dO/dtheta
 * divided by 1e-4 + O.
 */
#ifndef ORC_MLP_H
#define ORC_MLP_H

#include "sj.h"

#define ORC_MLP_NIN 13
#define ORC_MLP_HID 64
#define ORC_MLP_NOUT 4
#define ORC_MLP_NPARAM 9476
#define ORC_MLP_LS0 (-1.3862943611198906) /* ln 0.25 */
#define ORC_MLP_LS_HALF 3.0
#define ORC_MLP_REL_D0 0.1
#define ORC_MLP_LOG_EPS 1e-4

enum { ORC_MLP_W1 = 0, ORC_MLP_B1 = 832, ORC_MLP_W2 = 896, ORC_MLP_B2 = 4992, ORC_MLP_W3 = 5056,
       ORC_MLP_B3 = 9152, ORC_MLP_W4 = 9216, ORC_MLP_B4 = 9472 };

void orc_mlp_features(const orc_sj *wf, const double *cfg, int i, const double *r, double *f);

/* bf16 round-to-nearest-even of a float (the device's operand rounding) */
float orc_bf16(float x);

/* out[4] = network(f); with emulate_bf16 the weights, inputs and hidden
 * activations are rounded to bf16 as the MFMA kernels round them (sums and
 * tanh in FP64).  act (optional) receives the three hidden layers. */
void orc_mlp_forward(const float *w, const double *f, int emulate_bf16, double *out, double *act);

/* ls from the raw output (smoothly bounded) and its slope d ls / d out3 */
double orc_mlp_log_sigma(double out3);
double orc_mlp_log_sigma_slope(double out3);

/* One training sample: electron i of cfg proposes with normals z.  Returns
 * the objective (0, 1 or 2 as above) of the sample; adds its derivative to
 * grad (if non-null); diag (if non-null)
 * receives {log r, A, |dx|^2, ls, ls', lR, x'[3]}.  frozen_f2 (optional)
 * replaces the reverse-evaluation inputs (finite-difference checks). */
double orc_mlp_sample(const orc_sj *wf, const float *w, const double *cfg, int i, const double *z,
                      int emulate_bf16, const double *frozen_f2, double *grad, double *diag, int objective);

/* log q(x|x') - log q(x'|x) of the proposal x -> x' of electron i */
double orc_mlp_log_q_ratio(const orc_sj *wf, const float *w, const double *cfg, int i, const double *xn,
                           int emulate_bf16);

#endif
