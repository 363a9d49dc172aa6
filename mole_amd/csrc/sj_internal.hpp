// Internal structures of the Slater-Jastrow walker path.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/mole_gpu.h"
#include "mlp.hpp"

namespace mole {

// At most six parameters: b_anti, b_para, c, a, d, and s (the basis
// exponents are s zeta0_m), the sixth only when the system declares it.
constexpr int kSJNParam = 6;

// Proposals of the Slater-Jastrow kernel and their draws per move.
enum SJProposal : int { kSJPropBox = 0, kSJPropDiffuse = 1, kSJPropLearned = 2 };
__host__ __device__ constexpr int sj_draws_per_move(int prop) { return prop == kSJPropBox ? 4 : 6; }
constexpr int kSJMaxElectrons = 40;
constexpr int kSJMaxSpin = 20;

// Per-walker accumulator columns of Slater-Jastrow systems, [column][W].
enum SJAccColumn : int {
  kSJSumE = 0,
  kSJSumE2 = 1,
  kSJSumBm = 2,
  kSJSumBm2 = 3,
  kSJBlockSum = 4,
  kSJAccepted = 5,
  kSJCondMax = 6,    // largest ||A||_1 ||A^-1||_1 seen at an LU refresh
  kSJLdDrift = 7,    // largest |log|det| accumulated - log|det| refreshed| (FP64)
  kSJSingular = 8,   // spin blocks found singular at a refresh
  kSJSumO = 9,  // then sum O [np], sum O E [np], sum O O^T [np (np + 1) / 2]
};
// Column layout of a system with np parameters (5, or 6 with the exponent
// scale): only the system's own parameters take columns, in HBM and in LDS.
__host__ __device__ constexpr int sj_sum_oe(int np) { return kSJSumO + np; }
__host__ __device__ constexpr int sj_sum_oo(int np) { return kSJSumO + 2 * np; }  // upper triangle, row-major
__host__ __device__ constexpr int sj_columns(int np) { return kSJSumO + 2 * np + np * (np + 1) / 2; }
constexpr int kSJColumnsMax = sj_columns(kSJNParam);
// The columns reduced by their maximum (the rest by their sum).
constexpr int kSJMaxFirst = kSJCondMax, kSJMaxCount = 2;

// Per-walker log-determinant state, [kSJLdColumns][W]: for each spin block
// (mantissa, exponent) with det A_s = mantissa 2^exponent, the mantissa in
// [1/2, 1) in magnitude (0: singular).
constexpr int kSJLdColumns = 4;

// One system as the kernels see it (device memory).
struct SJSysDev {
  int N, n_up, n_dn, n_nuc, n_basis, n_orb;
  int L;  // lanes per walker (power of two, <= 64)
  int hamiltonian;
  int W;
  const double *nuc;     // [n_nuc][3]
  const double *charge;  // [n_nuc]
  const double *zeta;    // [n_basis]
  const int *center;     // [n_basis]
  const int *kind;       // [n_basis] MOLE_BASIS_*
  int any_gto;           // 0: no Gaussian basis function
  int any_p;             // 0: no p-type STO (with any_gto == 0, kind is not read)
  int opt_exp;           // 1: p[5] = s is a parameter (zeta holds s zeta0)
  const double *coef_t;  // [n_basis][n_orb]: lanes over orbitals read consecutive words
  double p[kSJNParam];
  double v_nn;
  double *pos;          // [3N][W]
  void *B;              // [n_up^2 + n_dn^2][W], element type of the launch
  double *acc;          // [sj_columns(5 + opt_exp)][W]
  double *ld;           // [kSJLdColumns][W] log-determinant state
  uint64_t *trace_mask; // [W][cap_sweeps] or null
  double *trace_e;      // [W][cap_samples] or null
  // compat samples (extended instantiation): psi_s [W][cap_samples] and
  // O psi_s [W][cap_samples][np], or null
  double *samp_psi, *samp_g;
  uint64_t draw_base;   // draw index of the first move after the stream origin
  uint64_t move_base;   // moves made before the stream origin (LU refresh phase)
  uint32_t sys_index;   // mixed into the Philox key
  uint32_t lds_per_walker;  // bytes
  // diffusion Monte Carlo (null outside DMC): [4][W] = weight, local
  // energy at the current configuration, and the last step's w E_old and
  // w_old; the reference energy is a device scalar
  double *dmc;
  const double *dmc_eref;
  // the walkers a DMC step starts from, as the branching gathered them:
  // positions, inverses, cached local energies and weights (the step writes
  // the moved walkers to pos, B and dmc)
  const double *dmc_pos_in;
  const void *dmc_B_in;
  const double *dmc_el_in, *dmc_w_in;
  // local operators: ops[0] the custom Hamiltonian (when custom_h), ops[1..n_obs]
  // the observables, whose per-walker sums are obs_acc [2 n_obs][W] (sum, sum of squares)
  const mole_operator *ops;
  int custom_h, n_obs;
  double *obs_acc;
  // kSJEval outputs per walker (each optional): psi [W], d psi [W][N][3],
  // lap psi [W], E_L [W], d psi/dp [W][np]
  double *ev_value, *ev_grad, *ev_lap, *ev_el, *ev_pgrad;
};

// What a launch does besides the sweeps: nothing (VMC), DMC's start (local
// energies and unit weights of the configurations), or DMC steps.
// kSJEval (mole_eval, extended kernel, no sweeps): the value, gradient,
// laplacian, local energy and parameter gradient of each walker's
// configuration into the ev_* arrays.
enum SJDmcMode : int { kSJDmcOff = 0, kSJDmcInit = 1, kSJDmcStep = 2, kSJEval = 3 };

// A workgroup's work: walkers [w0, w0 + nw) of one system.
struct SJTile {
  int sys, w0, nw;
};

// Launch state shared by every walker of a context, in device memory so a
// captured sequence [sweep kernel, advance kernel] can be replayed: the
// sweep kernel reads it, the advance kernel moves it on.
struct SJLaunchState {
  uint64_t sweeps_done;   // sweeps since the stream origin
  uint32_t block_fill;    // samples in the block in progress
  uint32_t trace_sweeps;  // trace cursors since the last reset
  uint32_t trace_samples;
  uint32_t pad;
};

struct SJArgs {
  SJLaunchState *state;
  const SJSysDev *systems;
  const SJTile *tiles;
  uint64_t seed;
  uint32_t epoch;
  double scale, offset;    // Range(-side/2, side/2)
  double tau, sqrt_tau;    // drift-diffusion time step
  uint32_t n_sweeps;
  uint32_t sample_start;   // sweeps s >= sample_start are sampled
  uint32_t block_size;
  uint32_t lu_period;      // 0: never
  uint32_t cap_sweeps, cap_samples;
  int init;                // 1: draw U(-1,1) configurations first (draws 0..3N-1)
  int refresh;             // 1: rebuild the inverses from the positions at entry;
                           // 2: only the log-determinant state (inverses kept)
  MlpDev net;              // learned proposal: the packed network
  int dmc;                 // SJDmcMode
  double dmc_tau;          // DMC: the weights' time step
  int ext;                 // 1: DMC, local operators or Gaussian bases in use (the extended kernel)
  int one_wave;            // learned proposal: the kernel fitted to one wave per SIMD
  uint32_t *ticket;        // persistent launch: next tile index (zero on entry)
  int n_tiles;             // persistent launch: tiles in `tiles`
};

size_t sj_lds_per_walker(int N, int nu, int nd, int nn, int K, int tsize, int np);

// Draws start at draw_base + D N sweeps_done (D draws per move of the
// proposal) and moves at move_base + N sweeps_done for every system; sample
// s of a launch is sampled when s >= sample_start.
hipError_t launch_sj(const SJArgs &a, int n_tiles, int fp32, int proposal, size_t lds_bytes,
                     hipStream_t stream);
// One persistent launch of the box-proposal plain kernel over a.n_tiles tiles
// (a.tiles in the order they are taken), then the ticket's reset.
hipError_t launch_sj_persistent(const SJArgs &a, int workgroups, int fp32, size_t lds_bytes, hipStream_t stream);
// Resident workgroups of the persistent kernel per CU at lds_bytes of LDS.
int sj_persistent_occupancy(int fp32, size_t lds_bytes);
// One device optimiser step of a Slater-Jastrow system from its reduced
// accumulator columns: params (n_params = 5, or 6 with the exponent scale)
// and sys->p updated in place, zeta = s zeta0 with the scale; log[8] as the
// two-electron step's.
hipError_t launch_sj_vmc_step(const double *red, double n_samples, int n_params, double *params, mole_optimizer *opt,
                              double *log, SJSysDev *sys, const double *zeta0, double *zeta, int n_basis,
                              hipStream_t stream);
// state->sweeps_done += n_sweeps; block fill and trace cursors advance by the
// launch's samples.
hipError_t launch_sj_advance(SJLaunchState *state, uint32_t n_sweeps, uint32_t n_samples,
                             uint32_t block_size, hipStream_t stream);
// *state = v, by a kernel (no memset nodes if the stream is being captured)
hipError_t launch_sj_state_set(SJLaunchState *state, const SJLaunchState &v, hipStream_t stream);

}  // namespace mole
