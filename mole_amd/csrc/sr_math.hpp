// The optimizer step from reduced sums, shared by the host entry point
// mole_compute_parameter_update and the device-side optimisation loop so
// that both round identically: SteepestDescent (src/optimize/src/
// optimizers.rs:20-30) and StochasticReconfiguration (optimizers.rs:181-253)
// with the energy gradient of src/optimize/src/util.rs:6-46.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mole_gpu.h"

namespace mole {

// Returns 0, or MOLE_ERR_LINALG for a singular SR matrix.  g (optional)
// receives the energy gradient.
__host__ __device__ inline int sr_update(int np, double n, double sum_e, const double *sum_o, const double *sum_oe,
                                         const double *sum_oo, uint32_t optimizer, double step, uint32_t sr_mode,
                                         double sr_eps, double *dp, double *g_out) {
  const double ebar = sum_e / n;
  double g[MOLE_MAX_PARAMS], obar[MOLE_MAX_PARAMS];
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
  double A[MOLE_MAX_PARAMS][MOLE_MAX_PARAMS + 1];
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

}  // namespace mole
