// Host-visible launch interface of the walker-MLP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"
#include "sj_internal.hpp"

namespace mole {

struct MlpTrainArgs {
  MlpDev net;
  const SJSysDev *sys;  // the system the configurations belong to
  const double *cfg;    // [n_cfg][3N] positions
  uint32_t n_cfg;
  uint64_t first;       // global index of the batch's first sample
  uint32_t batch;       // samples, a multiple of 16
  uint64_t seed;        // Philox key of the proposal normals
  float *partial;       // [gridDim][kMlpParams], written by the kernel (no clearing needed)
  double *obj;          // [gridDim] sum of the objective, written by the kernel
  int objective;        // MOLE_MLP_OBJ_*: ESJD; RELATIVE, each sample's
                        // A |dx|^2 divided by d^2 + kMlpRelD0^2 (d the larger
                        // nearest-nucleus distance of x and x'); LOG, each
                        // sample's log(kMlpLogEps + A |dx|^2)
};
// length scale below which the relative objective stops growing (bohr)
constexpr double kMlpRelD0 = 0.1;
// floor of the log objective: a rejected move scores log(1e-4), a jump of
// 0.01 bohr
constexpr double kMlpLogEps = 1e-4;

// Philox system tag of the training normals: sample g draws 0..3 of the
// stream (seed, kMlpTrainSys, lo32(g), hi32(g)).
constexpr uint32_t kMlpTrainSys = 0x6D6C70u;
// Philox epoch of the Xavier initialiser (walker 0, draw p = parameter p).
constexpr uint32_t kMlpInitEpoch = 0x58415631u;

hipError_t launch_mlp_init(float *w, uint64_t seed, hipStream_t stream);
hipError_t launch_mlp_pack(const float *w, bf16x4 *fwd1, bf16x8 *fwd, bf16x4 *bwd, float *bias,
                           hipStream_t stream);
hipError_t launch_mlp_train(const MlpTrainArgs &a, int n_groups, hipStream_t stream);
// One Adam step on the loss -mean(A |dx|^2) (g = -grad / batch) of the FP32
// master weights; step counts from 1 (bias correction)
struct MlpAdamStep {
  float *w, *m, *v;
  float inv_batch, lr, beta1, beta2, eps;
  int step;
};
// grad[p] = sum_g partial[g][p], obj_sum = sum_g obj[g] (fixed order: with
// more than kMlpSlice groups, sums of consecutive slices of kMlpSlice rows
// into mid [ceil(G / kMlpSlice)][kMlpParams] and mid_obj first, then those in
// slice order); with adam, the same thread then applies the Adam step to
// parameter p
constexpr int kMlpSlice = 32;
hipError_t launch_mlp_reduce(const float *partial, const double *obj, int n_groups, float *mid, double *mid_obj,
                             float *grad, double *obj_sum, const MlpAdamStep *adam, hipStream_t stream);
// features and network outputs of given configurations (debug / tests)
hipError_t launch_mlp_eval(const MlpDev &net, const double *cfg, int N, const double *nuc, int n_nuc,
                           const int *electron, int n, float *out, hipStream_t stream);

}  // namespace mole
