// Internal structures shared by the host side of the C ABI and the kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace mole {

constexpr int kMaxNuc2 = 8;   // nuclei of the two-electron kinds
constexpr int kNParam2 = 3;   // (zeta, b, c)

// Per-walker accumulator columns, stored [column][walker].
enum AccColumn : int {
  kAccSumE = 0,
  kAccSumE2 = 1,
  kAccSumBm = 2,
  kAccSumBm2 = 3,
  kAccBlockSum = 4,   // partial sum of the block in progress
  kAccAccepted = 5,   // accepted moves (exact integers in a double)
  kAccSumO = 6,       // kNParam2 columns
  kAccSumOE = 6 + kNParam2,
  kAccSumOO = 6 + 2 * kNParam2,  // kNParam2*(kNParam2+1)/2 columns, upper triangle
  kAccColumns2 = 6 + 2 * kNParam2 + kNParam2 * (kNParam2 + 1) / 2
};

struct Sys2 {
  int kind;         // MOLE_WF_HELIUM_PADE or MOLE_WF_LCAO2_PADE
  int hamiltonian;  // MOLE_H_*
  int n_nuc;
  double nuc[kMaxNuc2][3];
  double charge[kMaxNuc2];
  double v_nn;      // nuclear repulsion, summed as IonicPotential::new does
  double zeta, b, c;
};

struct Sweep2Args {
  Sys2 sp;
  const double *dparams;  // if set, (zeta, b, c) are read from device memory
  int W;             // walkers of this system
  double *pos;       // [6][W]
  double *acc;       // [kAccColumns2][W]
  int fresh;         // 1: every walker's accumulators start from zero (as after a
                     // reset) instead of being read
  // stream
  uint64_t seed;
  uint32_t sys, epoch;
  uint64_t draw0;    // draw index of the first proposal of this launch
  // box proposal Range(lo, hi) as scale/offset
  double scale, offset;
  uint32_t n_sweeps;
  uint32_t sample_start;  // sweeps s >= sample_start are sampled
  uint32_t block_size;
  uint32_t block_fill;    // samples already in the block in progress
  int o_conv;
  // trace (optional)
  uint64_t *trace_mask;   // [W][cap_sweeps]
  double *trace_e;        // [W][cap_samples]
  uint32_t trace_sweep0, trace_sample0, cap_sweeps, cap_samples;
  double *samp_psi;       // [W][cap_samples]
  double *samp_g;         // [W][cap_samples][3]
};

struct Eval2Args {
  Sys2 sp;
  size_t n;
  const double *cfg;  // [n][6]
  double *value, *grad, *lap, *e_local, *pgrad;
};

// variant: 0 auto, 1 one lane per walker, 2 wave-specialised (when eligible)
bool sweep2_wavespec_eligible(const Sys2 &sp);

// One optimizer step on the device from the reduced accumulator columns
// (red[kAccColumns2]) of n_samples samples: params += dp (same rounding as
// mole_compute_parameter_update); log[0..7] = sum_e, sum_e2, sum of block
// means, sum of squared block means, accepted moves, status, n_samples, 0.
// per block b in [first, first + k): (sum over walkers of block means, sum
// of their squares) from the E_L trace [W][cap], blocks of B samples
hipError_t launch_block_records(const double *trace_e, int W, size_t cap, size_t B, size_t first, size_t k,
                                double *out, hipStream_t stream);
hipError_t launch_vmc_step(const double *red, double n_samples, double *params, mole_optimizer *opt, double *log,
                           hipStream_t stream);
hipError_t launch_sweep2(const Sweep2Args &a, int variant, int block_threads,
                         hipStream_t stream);
hipError_t launch_init2(double *pos, int W, uint64_t seed, uint32_t sys,
                        uint32_t epoch, hipStream_t stream);
hipError_t launch_eval2(const Eval2Args &a, hipStream_t stream);
// out[k] = sum over walkers of column k of acc [K][W] in a fixed order; the
// columns [max_first, max_first + max_count) by their maximum instead
hipError_t launch_reduce_columns(const double *acc, int W, int K, double *out,
                                 int max_first, int max_count, hipStream_t stream);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t stream);
// tau_out[w] = integrated autocorrelation time (scripts/statfor.rs:30-54) of
// samples [first, first + n) of walker w's series in trace_e [W][cap], n >= 2
hipError_t launch_tau_corr(const double *trace_e, int W, size_t cap, size_t first, int n, double *tau_out,
                           hipStream_t stream);

}  // namespace mole
