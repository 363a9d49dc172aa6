// The optimizer step from reduced sums, shared by the host entry points
// (mole_compute_parameter_update, mole_optimizer_update) and the
// device-side optimisation loop so that all of them round identically:
// SteepestDescent (src/optimize/src/optimizers.rs:20-30),
// StochasticReconfiguration (optimizers.rs:181-253), MomentumDescent
// (:32-61), NesterovMomentum (:63-95) and OnlineLbfgs (:97-179), with the
// energy gradient of src/optimize/src/util.rs:6-46.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mole_gpu.h"

namespace mole {

// Returns 0, or MOLE_ERR_LINALG for a singular SR matrix.  g (optional)
// receives the energy gradient.  MAXP sizes the working arrays; below
// MOLE_MAX_PARAMS it is also the parameter count (a caller with a fixed
// count, whose small arrays then stay in registers), with the same
// arithmetic in the same order.
template <int MAXP = MOLE_MAX_PARAMS>
__host__ __device__ inline int sr_update(int np, double n, double sum_e, const double *sum_o, const double *sum_oe,
                                         const double *sum_oo, uint32_t optimizer, double step, uint32_t sr_mode,
                                         double sr_eps, double *dp, double *g_out) {
  if (MAXP < MOLE_MAX_PARAMS) np = MAXP;
  const double ebar = sum_e / n;
  double g[MAXP], obar[MAXP];
  for (int i = 0; i < np; ++i) {
    obar[i] = sum_o[i] / n;
    g[i] = 2.0 * (sum_oe[i] - ebar * sum_o[i]) / n;
  }
  if (g_out)
    for (int i = 0; i < np; ++i) g_out[i] = g[i];
  if (optimizer == MOLE_OPT_SD) {
    for (int i = 0; i < np; ++i) dp[i] = -(step * g[i]);
    return 0;
  }
  // augmented matrix [S | -g/2], Gaussian elimination with partial pivoting
  double A[MAXP][MAXP + 1];
  double osum = 0.0;
  for (int i = 0; i < np; ++i) osum += obar[i];
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      double s = sum_oo[i * np + j] / n;
      s -= (sr_mode == MOLE_SR_REFERENCE) ? osum * osum : obar[i] * obar[j];
      A[i][j] = s;
    }
    A[i][i] *= 1.0 + ((sr_mode == MOLE_SR_REFERENCE) ? 1e-2 : sr_eps);
    A[i][np] = -0.5 * g[i];
  }
  for (int k = 0; k < np; ++k) {
    int piv = k;
    for (int i = k + 1; i < np; ++i)
      if (fabs(A[i][k]) > fabs(A[piv][k])) piv = i;
    if (!(fabs(A[piv][k]) > 0.0)) return MOLE_ERR_LINALG;
    if (piv != k)
      for (int j = 0; j <= np; ++j) {
        const double t = A[k][j];
        A[k][j] = A[piv][j];
        A[piv][j] = t;
      }
    for (int i = k + 1; i < np; ++i) {
      const double f = A[i][k] / A[k][k];
      for (int j = k; j <= np; ++j) A[i][j] -= f * A[k][j];
    }
  }
  for (int i = np - 1; i >= 0; --i) {
    double s = A[i][np];
    for (int j = i + 1; j < np; ++j) s -= A[i][j] * dp[j];
    dp[i] = s / A[i][i];
  }
  for (int i = 0; i < np; ++i) dp[i] = step * dp[i];
  return 0;
}

__host__ __device__ inline double dot_n(int n, const double *a, const double *b) {
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += a[i] * b[i];
  return t;
}

// OnlineLbfgs::update_curvature_pairs (optimizers.rs:146-158): push
// (pars - pars_prev, grad - grad_prev), dropping the oldest pair when
// `history` are held.
__host__ __device__ inline void lbfgs_push(mole_optimizer *o, const double *pars, const double *g) {
  const int np = (int)o->n_params;
  uint32_t slot;
  if (o->n_pairs >= o->history) {
    slot = o->head;
    o->head = (o->head + 1) % o->history;
  } else {
    slot = (o->head + o->n_pairs) % o->history;
    ++o->n_pairs;
  }
  for (int i = 0; i < np; ++i) {
    o->s[slot][i] = pars[i] - o->pars_prev[i];
    o->y[slot][i] = g[i] - o->grad_prev[i];
  }
}

// OnlineLbfgs::initial_direction (optimizers.rs:121-144): two-loop
// recursion over the held pairs (newest first, then oldest first).
__host__ __device__ inline void lbfgs_direction(const mole_optimizer *o, const double *g, double *p) {
  const int np = (int)o->n_params;
  const int m = (int)o->n_pairs;
  double alpha[2 * MOLE_LBFGS_MAX_HISTORY];
  for (int i = 0; i < np; ++i) p[i] = -g[i];
  for (int k = m - 1; k >= 0; --k) {
    const uint32_t j = (o->head + k) % o->history;
    const double a = dot_n(np, o->s[j], p) / dot_n(np, o->s[j], o->y[j]);
    for (int i = 0; i < np; ++i) p[i] = p[i] - a * o->y[j][i];
    alpha[k] = a;
  }
  double scale;
  if (o->iter == 0) {
    scale = 1e-10;
  } else {
    double tot = 0.0;
    for (int k = 0; k < m; ++k) {
      const uint32_t j = (o->head + k) % o->history;
      tot = tot + dot_n(np, o->s[j], o->y[j]) / dot_n(np, o->y[j], o->y[j]);
    }
    const uint64_t d = o->iter < o->history ? o->iter : (uint64_t)o->history;
    scale = tot / (double)d;
  }
  for (int i = 0; i < np; ++i) p[i] = p[i] * scale;
  for (int k = 0; k < m; ++k) {
    const uint32_t j = (o->head + k) % o->history;
    const double c = alpha[k] - dot_n(np, o->y[j], p) / dot_n(np, o->y[j], o->s[j]);
    for (int i = 0; i < np; ++i) p[i] = p[i] + c * o->s[j][i];
  }
}

// Optimizer::compute_parameter_update for every kind from the sums.
// Returns 0 or MOLE_ERR_LINALG (SR, state untouched).
// MAXP as for sr_update (below MOLE_MAX_PARAMS: o->n_params must equal it).
template <int MAXP = MOLE_MAX_PARAMS>
__host__ __device__ inline int optimizer_update(mole_optimizer *o, double n, double sum_e, const double *sum_o,
                                                const double *sum_oe, const double *sum_oo, const double *pars,
                                                double *dp, double *g_out) {
  const int np = MAXP < MOLE_MAX_PARAMS ? MAXP : (int)o->n_params;
  double g[MAXP];
  if (o->kind == MOLE_OPT_SD || o->kind == MOLE_OPT_SR) {
    const int rc =
        sr_update<MAXP>(np, n, sum_e, sum_o, sum_oe, sum_oo, o->kind, o->step, o->sr_mode, o->sr_eps, dp, g);
    if (g_out)
      for (int i = 0; i < np; ++i) g_out[i] = g[i];
    if (rc == 0) ++o->iter;
    return rc;
  }
  sr_update<MAXP>(np, n, sum_e, sum_o, sum_oe, sum_oo, MOLE_OPT_SD, 0.0, 0, 0.0, dp, g);
  if (g_out)
    for (int i = 0; i < np; ++i) g_out[i] = g[i];
  const double mu = o->momentum;
  if (o->kind == MOLE_OPT_MOMENTUM) {
    for (int i = 0; i < np; ++i) {
      o->m[i] = o->m[i] - o->step * g[i];
      dp[i] = mu * o->m[i];
    }
  } else if (o->kind == MOLE_OPT_NESTEROV) {
    for (int i = 0; i < np; ++i) {
      o->m_prev[i] = o->m[i];
      o->m[i] = mu * o->m[i] + o->step * g[i];
      dp[i] = -(mu * o->m_prev[i] + (1.0 + mu) * o->m[i]);
    }
  } else {  // MOLE_OPT_LBFGS
    double p[MAXP];
    lbfgs_push(o, pars, g);
    lbfgs_direction(o, g, p);
    for (int i = 0; i < np; ++i) dp[i] = -o->step * p[i];
    lbfgs_push(o, pars, g);
    for (int i = 0; i < np; ++i) {
      o->grad_prev[i] = g[i];
      o->pars_prev[i] = pars[i];
    }
  }
  ++o->iter;
  return 0;
}

}  // namespace mole
