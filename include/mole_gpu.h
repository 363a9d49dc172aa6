/*
 * mole_gpu.h -- C ABI of the MI355X (gfx950) walker hot path for mole.
 *
 * This is the boundary a Rust FFI crate binds: plain C types, caller-owned
 * host buffers valid for the duration of a call, no callbacks, no
 * exceptions.  A context owns its device buffers and one HIP stream; it is
 * not internally synchronised (one host thread at a time).
 *
 * Each entry point names the reference interface it replaces
 * (Jvanrhijn/mole, paths relative to the repository root).
 */
#ifndef MOLE_GPU_H
#define MOLE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------
 * 1..6 are the variants of errors::Error (src/errors/src/lib.rs:7-15) in
 * declaration order.  The remaining codes have no variant there; a Rust
 * binding maps them to Error::FuncError. */
typedef int32_t mole_status;
#define MOLE_OK 0
#define MOLE_ERR_LINALG 1                 /* Error::LinalgError */
#define MOLE_ERR_SHAPE 2                  /* Error::ShapeError */
#define MOLE_ERR_FUNC 3                   /* Error::FuncError */
#define MOLE_ERR_OPERATOR_VALUE_ACCESS 4  /* Error::OperatorValueAccessError */
#define MOLE_ERR_DATA_ACCESS 5            /* Error::DataAccessError */
#define MOLE_ERR_EMPTY_CACHE 6            /* Error::EmptyCacheError */
#define MOLE_ERR_INVALID_ARGUMENT 16      /* reference panics (asserts) */
#define MOLE_ERR_DEVICE 17                /* HIP runtime failure */
#define MOLE_ERR_UNSUPPORTED 18
#define MOLE_ERR_OUT_OF_MEMORY 19

/* ---- wavefunctions, Hamiltonians, proposals --------------------------- */
/* psi = exp(-zeta (r1 + r2) + b r12/(1 + c r12)) about nucleus 0,
 * params (zeta, b, c); one electron of each spin. */
#define MOLE_WF_HELIUM_PADE 1
/* psi = phi(r1) phi(r2) exp(b r12/(1 + c r12)),
 * phi(r) = sum_I exp(-zeta |r - R_I|), params (zeta, b, c). */
#define MOLE_WF_LCAO2_PADE 2
/* psi = det A_up det A_dn exp(J), A_s[c][k] = phi_k(r of the c-th electron
 * of spin s), phi_k(r) = sum_m coef[k][m] chi_m(r - R_center(m)) with basis
 * functions chi_m of exponent zeta_m and kind MOLE_BASIS_*,
 * J = sum_{i<j} b_ij r_ij/(1 + c r_ij) - sum_{i,I} a Z_I r_iI/(1 + d r_iI),
 * b_ij = b_anti (opposite spins) or b_para (same spin); params
 * (b_anti, b_para, c, a, d), or (b_anti, b_para, c, a, d, s) with the
 * exponent scale s as a sixth variational parameter: the basis functions
 * then take exponents s basis_exponent[m] (the orbital exponents the
 * reference's example wavefunctions optimise, e.g.
 * examples/helium_atom_singlet.rs:102-111, tests/sho_optimize.rs:90-106).
 * Electrons 0..n_up-1 are spin up.  A context holds either two-electron
 * kinds or Slater-Jastrow systems, not both. */
#define MOLE_WF_SLATER_JASTROW 3

/* Slater-Jastrow precision: FP64 throughout, or MIXED = Sherman-Morrison
 * inverses, ratios and Jastrow differences in FP32 with FP64 positions,
 * FP64 Gauss-Jordan re-inversion, FP64 local energy and sums. */
#define MOLE_PREC_FP64 0
#define MOLE_PREC_MIXED 1

#define MOLE_H_ELECTRONIC 0 /* ElectronicHamiltonian, src/operator/src/operator.rs:157-183 */
#define MOLE_H_IONIC 1      /* IonicHamiltonian, operator.rs:131-149 */

#define MOLE_PROPOSAL_BOX 0 /* MetropolisBox, src/metropolis/src/metrop.rs:21-101 */
/* MetropolisDiffuse, metrop.rs:103-212: x' = (x_i + tau v) + sqrt(tau) z with
 * v = grad_i psi / psi, a sign change rejects, Green's functions over the
 * moved electron's row (the reference's norms run over every row).  step is
 * tau.  Slater-Jastrow systems only; 6 draws per move (Box-Muller normals
 * from draws 0-3, accept uniform from draw 4). */
#define MOLE_PROPOSAL_DIFFUSE 1
/* Learned Metropolis-Hastings proposal (no counterpart in the reference; it
 * plugs into the propose/accept contract of metropolis/src/traits.rs:10-38):
 * an MLP 13 -> 64 -> 64 -> 64 -> 4 (tanh) maps electron i's neighbourhood to
 * a shift mu and ln sigma, x' = (x_i + mu) + sigma z, and the move is
 * accepted with the exact ratio min(R^2 q(x|x')/q(x'|x), 1).  step is
 * ignored; the network comes from mole_mlp_attach.  Slater-Jastrow systems,
 * 4 lanes per walker, 6 draws per move as MOLE_PROPOSAL_DIFFUSE. */
#define MOLE_PROPOSAL_LEARNED 2

/* Convention of the "Parameter gradient" observable O.
 * LOGDERIV:  O = d ln psi / dp.
 * REFERENCE: the reference's literal value, O = G/psi with the stored
 *            G = psi ./ (psi dpsi/dp) produced by Vector/Scalar
 *            (src/operator/src/traits.rs:150, src/vmc/src/operators.rs:9-13,
 *            src/optimize/src/util.rs:42). */
#define MOLE_O_LOGDERIV 0
#define MOLE_O_REFERENCE 1

/* run flags */
#define MOLE_FLAG_TRACE 1u   /* record per-sweep accept masks and E_L */
#define MOLE_FLAG_SAMPLES 2u /* keep per-sample E, psi, G (compat data) */

/* mole_sweep flags */
#define MOLE_SWEEP_SAMPLE 1u /* sample observables after every sweep */

#define MOLE_MAX_ELECTRONS 40
#define MOLE_MAX_NUCLEI 64
#define MOLE_MAX_PARAMS 16

typedef struct mole_ctx mole_ctx;

/* One system: nuclei, electrons, wavefunction and walker count.  The
 * nuclei and charges are those of IonicPotential::new
 * (operator.rs:40-55). */
typedef struct {
  uint32_t wf_kind;     /* MOLE_WF_* */
  uint32_t hamiltonian; /* MOLE_H_* */
  uint32_t n_up, n_down;
  uint32_t n_nuclei;
  const double *nuclei_xyz;     /* [n_nuclei][3] */
  const int32_t *nuclei_charge; /* [n_nuclei] */
  uint32_t n_params;
  const double *params; /* Optimize::parameters, src/optimize/src/traits.rs:13 */
  uint32_t n_walkers;
  /* MOLE_WF_SLATER_JASTROW only (ignored otherwise) */
  uint32_t n_basis;
  const double *basis_exponent;  /* [n_basis] zeta_m (times the exponent scale, if any) */
  const uint32_t *basis_center;  /* [n_basis] nucleus index */
  uint32_t n_orbitals;           /* >= max(n_up, n_down) */
  const double *orbital_coeff;   /* [n_orbitals][n_basis] */
  const uint32_t *basis_kind;    /* [n_basis] MOLE_BASIS_*; NULL: all STO */
} mole_system_desc;

/* Slater-Jastrow basis functions about their centre R: */
#define MOLE_BASIS_STO 0 /* exp(-zeta |r - R|) */
#define MOLE_BASIS_GTO 1 /* exp(-zeta |r - R|^2) (the Gaussian trial functions of
                            examples/custom_operator.rs:64-100, examples/dmc.rs:33-70) */
/* p-type STOs d_q exp(-zeta |d|), d = r - R, q = x, y, z: the angular
 * functions of the 2p shell (the 2p hydrogen eigenfunction is
 * z exp(-r/2), E = -1/8). */
#define MOLE_BASIS_STO_PX 2
#define MOLE_BASIS_STO_PY 3
#define MOLE_BASIS_STO_PZ 4

/* How the walkers move and what they record. */
typedef struct {
  uint64_t seed;         /* Philox key (lo32, hi32 ^ system index) */
  uint32_t proposal;     /* MOLE_PROPOSAL_* */
  double step;           /* MetropolisBox box_side (metrop.rs:26,66), or the
                            MetropolisDiffuse time step (metrop.rs:104-108) */
  uint32_t block_size;   /* Runner::run block_size (src/montecarlo/src/montecarlo.rs:24-46) */
  uint32_t o_convention; /* MOLE_O_* */
  uint32_t flags;        /* MOLE_FLAG_* */
  uint32_t record_sweeps;/* capacity, in sweeps, of trace/sample buffers */
  uint32_t precision;    /* MOLE_PREC_* (Slater-Jastrow) */
  uint32_t lu_period;    /* re-invert the Slater matrices after every
                            lu_period moves; 0 = never (Slater-Jastrow) */
} mole_run_desc;

/* Energy statistics since the last reset, summed over the system's
 * walkers.  Block means are over block_size consecutive samples of one
 * walker, as VmcRunner::process_monte_carlo_results chunks the
 * worker-ordered sample list (src/vmc/src/vmc.rs:150-170). */
typedef struct {
  uint64_t n_samples;
  uint64_t n_blocks;
  uint64_t n_moves;
  uint64_t n_accepted;
  double sum_e, sum_e2;
  double sum_block_mean, sum_block_mean2;
  /* Slater-Jastrow: largest ||A||_1 ||A^-1||_1 met at a re-inversion since
   * the last reset (0 otherwise) -- the condition estimate that says
   * whether FP32 inverses are trustworthy. */
  double max_condition;
  /* Slater-Jastrow: spin blocks found numerically singular at a
   * re-inversion (a pivot |p| <= n 2^-52 ||A||_1): the division by the pivot
   * is skipped and the block's log|det| is -inf until the next
   * re-inversion. */
  uint64_t n_singular;
  /* Slater-Jastrow VMC: largest |log|det A_s| accumulated over the
   * Sherman-Morrison updates since the previous re-inversion - log|det A_s|
   * from the re-inversion's pivots|, both FP64 (the accumulation multiplies
   * the ratios of the update's precision): the error the FP32 (MIXED) or
   * FP64 update chain has built up between re-inversions. */
  double max_logdet_drift;
} mole_stats;

/* Sums for the energy gradient and the SR matrix
 * (src/optimize/src/util.rs:6-46, src/optimize/src/optimizers.rs:191-233),
 * over the same samples as mole_stats.  sum_oo is the full n_params^2
 * matrix, row-major. */
typedef struct {
  uint32_t n_params;
  uint64_t n;
  double sum_e;
  double sum_o[MOLE_MAX_PARAMS];
  double sum_oe[MOLE_MAX_PARAMS];
  double sum_oo[MOLE_MAX_PARAMS * MOLE_MAX_PARAMS];
} mole_sr_sums;

/* ---- library ---------------------------------------------------------- */
const char *mole_status_string(mole_status s);
/* Message of the last failing call on this thread ("" if none). */
const char *mole_last_error(void);
/* Device name and compute-unit count of the current device. */
mole_status mole_device_info(char *name, size_t cap, int32_t *cu_count);

/* ---- context -----------------------------------------------------------
 * Replaces Sampler::new / MetropolisBox::from_rng
 * (src/montecarlo/src/samplers.rs:40-55, metrop.rs:34-36) for a whole
 * batch of walkers of n_sys systems. */
mole_status mole_ctx_create(const mole_system_desc *sys, uint32_t n_sys,
                            const mole_run_desc *run, mole_ctx **out);
void mole_ctx_destroy(mole_ctx *ctx);

/* Sampler::new's initial configuration: U(-1, 1) per coordinate from draws
 * 0..3N-1 of each walker's stream (samplers.rs:46); the stream position is
 * left at 3N. */
mole_status mole_init_walkers(mole_ctx *ctx);
/* Sampler::with_initial_configuration (samplers.rs:57-71) and checkpoint:
 * configurations as [walker][electron][xyz]. */
mole_status mole_set_configs(mole_ctx *ctx, uint32_t sys, const double *cfg,
                             size_t n_walkers);
mole_status mole_get_configs(mole_ctx *ctx, uint32_t sys, double *cfg,
                             size_t n_walkers);
/* Every walker of sys takes the configuration cfg [electron][xyz]; the
 * per-iteration clone of the master sampler (src/vmc/src/vmc.rs:56). */
mole_status mole_broadcast_config(mole_ctx *ctx, uint32_t sys,
                                  const double *cfg);
/* Metropolis::reseed_rng (src/metropolis/src/traits.rs:30-37): selects the
 * Philox epoch (counter word 3) and the next draw index of every walker. */
mole_status mole_set_stream(mole_ctx *ctx, uint32_t epoch, uint64_t position);
mole_status mole_get_stream(mole_ctx *ctx, uint32_t *epoch,
                            uint64_t *position);

/* Optimize::update_parameters / parameters (src/optimize/src/traits.rs:8-16). */
mole_status mole_set_params(mole_ctx *ctx, uint32_t sys, const double *p,
                            uint32_t n_params);
mole_status mole_get_params(mole_ctx *ctx, uint32_t sys, double *p,
                            uint32_t n_params);

/* ---- sampling ----------------------------------------------------------
 * n_sweeps calls of Sampler::move_state (samplers.rs:106-117: electrons
 * 0..N-1 in order, MetropolisBox::move_state metrop.rs:84-96), each
 * followed, with MOLE_SWEEP_SAMPLE, by Sampler::sample of the "Energy",
 * "Wavefunction value" and "Parameter gradient" observables
 * (samplers.rs:81-104) folded into the statistics. */
mole_status mole_sweep(mole_ctx *ctx, uint32_t n_sweeps, uint32_t flags);
/* Runner::run(steps, block_size) (montecarlo.rs:24-46): steps/block_size
 * blocks of block_size sweeps, observables sampled from the second block
 * on.  Fails with MOLE_ERR_INVALID_ARGUMENT when steps < 2*block_size
 * (the assert at montecarlo.rs:29) or block_size differs from the
 * context's. */
mole_status mole_run(mole_ctx *ctx, uint32_t steps, uint32_t block_size);
mole_status mole_reset_stats(mole_ctx *ctx);
mole_status mole_read_stats(mole_ctx *ctx, uint32_t sys, mole_stats *out);
mole_status mole_read_sr_sums(mole_ctx *ctx, uint32_t sys, mole_sr_sums *out);
/* Per-sweep accept masks (bit e: move of electron e accepted) as
 * [walker][sweep] and local energies of sampled sweeps as [walker][sample],
 * since the last reset (needs MOLE_FLAG_TRACE). */
mole_status mole_read_trace(mole_ctx *ctx, uint32_t sys, uint64_t *accept_mask,
                            double *e_local, size_t walkers, size_t sweeps,
                            size_t samples);
/* Per-block records of the energy, streamed from the device: the
 * ensemble form of what Runner::run hands Log::log once per block
 * (src/montecarlo/src/montecarlo.rs:31-42, src/montecarlo/src/traits.rs:
 * 43-47) and of the block means VmcRunner::process_monte_carlo_results
 * forms (src/vmc/src/vmc.rs:150-170).  Block b covers samples
 * [b*block_size, (b+1)*block_size) since the last reset; a walker's block
 * mean is its samples summed in order times 1/block_size (the accumulators'
 * arithmetic), summed over walkers in a fixed order. */
typedef struct {
  uint64_t block;          /* index since the last statistics reset */
  uint64_t n_walkers;
  double mean;             /* sum_block_mean / n_walkers */
  double sum_block_mean;
  double sum_block_mean2;
} mole_block_record;
/* Records of the complete blocks [first, first + k), k = min(cap, blocks
 * complete - first); *n = k.  out may be NULL to count.  Needs
 * MOLE_FLAG_TRACE (records up to record_sweeps samples). */
mole_status mole_read_blocks(mole_ctx *ctx, uint32_t sys, size_t first, mole_block_record *out, size_t cap,
                             size_t *n);
/* The integrated autocorrelation time of every walker's local-energy
 * series, computed on the device from the trace (the series stay there):
 * correlation() of scripts/statfor.rs:30-54 over samples [first, first + n)
 * since the last reset,
 *   tau = max(1, 1 + 2 sum_{i=1..min(200, n-1)} rho_i), the sum cut at the
 *   first negative rho_i, rho_i = sum_t (x_t - m)(x_{t+i} - m)/((n - i) var),
 *   m the mean and var the variance with n - 1 (statfor.rs:86-88);
 * every sum in sample order, as the reference folds it.  tau[n_walkers].
 * Effective samples = sum_w n / tau_w.  Needs MOLE_FLAG_TRACE, n >= 2. */
mole_status mole_tau_int(mole_ctx *ctx, uint32_t sys, size_t first, size_t n, double *tau, size_t n_walkers);
/* mole_tau_int without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc): the per-walker values are
 * staged in a library-owned device array of MOLE_TAU_NOALLOC_MAX_WALKERS
 * doubles per GPU (held until the process ends), so the call makes no
 * hipMalloc/hipFree.  Same arguments and results as mole_tau_int; a system
 * with more walkers returns MOLE_ERR_UNSUPPORTED.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other threads, other
 * contexts' streams) race. */
#define MOLE_TAU_NOALLOC_MAX_WALKERS 262144
mole_status mole_tau_int_noalloc(mole_ctx *ctx, uint32_t sys, size_t first, size_t n, double *tau,
                                 size_t n_walkers);
/* MonteCarloSampler::data (src/montecarlo/src/traits.rs:27) in compat
 * form: "Energy", "Wavefunction value" and "Parameter gradient" samples as
 * [walker][sample] (and [walker][sample][param], the system's n_params), the
 * order in which VmcRunner::concatenate_worker_data appends worker data
 * (vmc.rs:108-130).  Needs MOLE_FLAG_SAMPLES.  Slater-Jastrow systems: psi_s
 * = (psi psi)/psi with psi = sign exp(log|det up| + log|det dn| + J) from
 * the carried FP64 determinants, and O_k psi_s with O = d ln psi/dp. */
mole_status mole_read_samples(mole_ctx *ctx, uint32_t sys, double *e_local,
                              double *psi, double *pgrad, size_t walkers,
                              size_t samples);
/* Slater-Jastrow: log|det A_up|, log|det A_dn| and their signs per walker,
 * as [walker][2] each -- the determinant factor of Function::value
 * (src/wavefunction_traits/src/lib.rs:7-11; psi = sign_up sign_dn
 * exp(logdet_up + logdet_dn + J)).  FP64, carried through every accepted
 * Sherman-Morrison update as log|det| <- log|det| + log|R_D| (a mantissa
 * product and a binary exponent, so nothing over- or underflows) and
 * replaced at every re-inversion by the sum of log|pivot| of its
 * Gauss-Jordan elimination.  A singular block reads -inf with sign 0.
 * DMC steps do not carry it (the next VMC re-inversion restores it).
 * sign may be NULL. */
mole_status mole_read_logdet(mole_ctx *ctx, uint32_t sys, double *logdet, double *sign,
                             size_t n_walkers);

/* ---- single configurations ---------------------------------------------
 * Function::value, Differentiate::gradient/laplacian
 * (src/wavefunction_traits/src/lib.rs:7-24), Optimize::parameter_gradient
 * (src/optimize/src/traits.rs:9) and the local energy
 * ElectronicHamiltonian::act_on / psi (operator.rs:177-183) for n
 * configurations [n][N][3].  grad is d psi (not d ln psi) as
 * [n][N][3]; pgrad is d psi/dp as [n][n_params].  Output pointers may be
 * NULL.  Slater-Jastrow systems are evaluated in FP64 whatever the
 * context's precision, from a Gauss-Jordan re-inversion at each
 * configuration (psi = sign exp(log|det up| + log|det dn| + J)), by the
 * sweep kernel on walkers of their own: the context's walkers are not
 * touched. */
mole_status mole_eval(mole_ctx *ctx, uint32_t sys, size_t n, const double *cfg,
                      double *value, double *grad, double *lap,
                      double *e_local, double *pgrad);

/* ---- optimizer step from device-reduced sums -------------------------
 * Optimizer::compute_parameter_update (src/optimize/src/traits.rs:18-25)
 * for SteepestDescent (src/optimize/src/optimizers.rs:20-30) and
 * StochasticReconfiguration (optimizers.rs:181-253), evaluated from the
 * sums of mole_read_sr_sums instead of per-sample vectors:
 *   Ebar = sum_e/n, g = 2 (sum_oe - Ebar sum_o)/n   (util.rs:6-46)
 *   SD: dp = -step g
 *   SR: S = sum_oo/n - Obar Obar^T (TEXTBOOK) or - (sum_i Obar_i)^2 on
 *       every element (REFERENCE, optimizers.rs:219-224); diag *= 1 + eps
 *       (REFERENCE: eps = 1e-2, optimizers.rs:225-231);
 *       dp = step S^{-1} (-g/2)  (solveh_into, optimizers.rs:251).
 * g may be NULL.  A singular S returns MOLE_ERR_LINALG. */
#define MOLE_OPT_SD 0
#define MOLE_OPT_SR 1
#define MOLE_SR_TEXTBOOK 0
#define MOLE_SR_REFERENCE 1
mole_status mole_compute_parameter_update(const mole_sr_sums *sums, uint32_t optimizer,
                                          double step, uint32_t sr_mode, double sr_eps,
                                          double *dp, double *g);

/* VmcRunner::run_optimization (src/vmc/src/vmc.rs:43-106) with persistent
 * walkers, run on the device: per iteration the statistics restart,
 * Runner::run(steps, block_size) sweeps the walkers, the accumulators are
 * reduced and the optimizer step (as mole_compute_parameter_update, same
 * rounding) updates the parameters in device memory -- no host round trip
 * between iterations.  log (optional, [iters][8]): sum_e, sum_e2, sum of
 * block means, sum of squared block means, accepted moves, status,
 * n_samples, 0.  params_out (optional): the final parameters, which the
 * context keeps.  One system (two-electron, or Slater-Jastrow with its 5
 * or 6 parameters: the step rewrites the device system table the sweep
 * kernel reads, and with the exponent scale the exponents s zeta0_m), no
 * trace/sample flags.  A singular SR matrix leaves the parameters unchanged
 * at that iteration and returns MOLE_ERR_LINALG after the loop. */
mole_status mole_optimize(mole_ctx *ctx, uint32_t iters, uint32_t steps, uint32_t block_size,
                          uint32_t optimizer, double step, uint32_t sr_mode, double sr_eps,
                          double *log, double *params_out);

/* ---- stateful optimizers ------------------------------------------------
 * MomentumDescent (src/optimize/src/optimizers.rs:32-61), NesterovMomentum
 * (:63-95) and OnlineLbfgs (:97-179) keep state between calls of
 * Optimizer::compute_parameter_update (src/optimize/src/traits.rs:18-25).
 * mole_optimizer is that state as plain data (it is copied to and from the
 * device by mole_optimize_with); SD and SR are accepted too and keep none.
 * With g the energy gradient from the sums (as above) and pars the current
 * parameters:
 *   MOMENTUM: m -= step g; dp = mu m                       (:49-60)
 *   NESTEROV: m_prev = m; m = mu m + step g;
 *             dp = -(mu m_prev + (1 + mu) m)                (:83-94)
 *   LBFGS:    the pair (pars - pars_prev, g - grad_prev) is pushed twice
 *             per call (:163,170), the oldest pairs dropped beyond
 *             `history`; two-loop recursion from -g scaled by 1e-10 on the
 *             first call, else by sum(s.y / y.y) / min(iter, history);
 *             dp = -step p; pars_prev and grad_prev start at 1e-5.
 * Dot products run in index order. */
#define MOLE_OPT_MOMENTUM 2
#define MOLE_OPT_NESTEROV 3
#define MOLE_OPT_LBFGS 4
#define MOLE_LBFGS_MAX_HISTORY 16
typedef struct {
  uint32_t kind;      /* MOLE_OPT_* */
  uint32_t n_params;
  uint32_t sr_mode;   /* SR only */
  uint32_t history;   /* LBFGS: pairs kept, 1..MOLE_LBFGS_MAX_HISTORY */
  double step;
  double sr_eps;      /* SR only */
  double momentum;    /* MOMENTUM / NESTEROV: momentum_parameter */
  uint64_t iter;      /* calls so far */
  uint32_t n_pairs;   /* LBFGS pairs held, oldest at index `head` */
  uint32_t head;
  double m[MOLE_MAX_PARAMS];
  double m_prev[MOLE_MAX_PARAMS];
  double grad_prev[MOLE_MAX_PARAMS];
  double pars_prev[MOLE_MAX_PARAMS];
  double s[MOLE_LBFGS_MAX_HISTORY][MOLE_MAX_PARAMS];
  double y[MOLE_LBFGS_MAX_HISTORY][MOLE_MAX_PARAMS];
} mole_optimizer;
/* The constructors MomentumDescent::new(step, mu, nparm) (:40-46),
 * NesterovMomentum::new (:71-78), OnlineLbfgs::new(step, history, nparm)
 * (:108-119), SteepestDescent::new(step) (:14-16) and
 * StochasticReconfiguration::new(step) (:187-189; sr_mode/sr_eps as for
 * mole_compute_parameter_update).  Unused arguments are ignored. */
mole_status mole_optimizer_init(mole_optimizer *opt, uint32_t kind, uint32_t n_params, double step,
                                double momentum, uint32_t history, uint32_t sr_mode, double sr_eps);
/* Optimizer::compute_parameter_update from device-reduced sums; pars
 * (n_params, read by LBFGS only) may be NULL for the other kinds; g may be
 * NULL.  A singular SR matrix returns MOLE_ERR_LINALG and leaves opt as it
 * was. */
mole_status mole_optimizer_update(mole_optimizer *opt, const mole_sr_sums *sums, const double *pars,
                                  double *dp, double *g);
/* mole_optimize with any optimizer: opt is copied to the device, advanced
 * once per iteration there, and copied back (its state after the loop). */
mole_status mole_optimize_with(mole_ctx *ctx, uint32_t iters, uint32_t steps, uint32_t block_size,
                               mole_optimizer *opt, double *log, double *params_out);

/* ---- local operators on the device --------------------------------------
 * LocalOperator<T>::act_on(wf, cfg) / psi (src/operator/src/traits.rs:
 * 263-265, as Sampler::sample divides, src/montecarlo/src/samplers.rs:81-104)
 * for operators built from these terms, evaluated per walker on the
 * device from the same derivatives as the local energy (Slater-Jastrow
 * systems):
 *   KINETIC               -1/2 lap(psi)/psi       KineticEnergy, operator.rs:96-121
 *   IONIC_POTENTIAL       -sum_{I,i} Z_I/|r_i - R_I| + ion-ion repulsion
 *                                                  IonicPotential, operator.rs:12-64
 *   ELECTRONIC_POTENTIAL  sum_{i<j} 1/r_ij        ElectronicPotential, operator.rs:66-94
 *   HARMONIC              1/2 w^2 sum_i |r_i - c|^2, w = param[0], c = param[1..3]
 *                                                  (the potential of examples/custom_operator.rs:48-61)
 *   RADIAL_POWER          sum_i |r_i - c|^k, k = param[0] in -4..4, k != 0, c = param[1..3]
 *   DIPOLE                sum_i (r_i - c).u, c = param[0..2], u = param[3..5]
 * An operator is sum_k coef_k term_k (at most MOLE_OPERATOR_MAX_TERMS terms),
 * e.g. HarmonicHamiltonian = KINETIC + HARMONIC, ElectronicHamiltonian =
 * KINETIC + IONIC_POTENTIAL + ELECTRONIC_POTENTIAL. */
#define MOLE_TERM_KINETIC 0
#define MOLE_TERM_IONIC_POTENTIAL 1
#define MOLE_TERM_ELECTRONIC_POTENTIAL 2
#define MOLE_TERM_HARMONIC 3
#define MOLE_TERM_RADIAL_POWER 4
#define MOLE_TERM_DIPOLE 5
#define MOLE_OPERATOR_MAX_TERMS 4
#define MOLE_MAX_OBSERVABLES 8
typedef struct {
  uint32_t kind; /* MOLE_TERM_* */
  double coef;
  double param[6];
} mole_term;
typedef struct {
  uint32_t n_terms;
  mole_term term[MOLE_OPERATOR_MAX_TERMS];
} mole_operator;
/* The operator whose local values are the system's "Energy": the energy
 * statistics, the SR sums, traces, block records and DMC weights use it
 * instead of the descriptor's hamiltonian (the operators! "Energy" entry of
 * examples/custom_operator.rs:124-126).  NULL restores the built-in one. */
mole_status mole_set_hamiltonian(mole_ctx *ctx, uint32_t sys, const mole_operator *h);
/* Further observables, sampled with the energy after every sampled sweep
 * (VMC): per walker sum O and sum O^2 since the last reset.  n = 0 clears
 * them. */
mole_status mole_set_observables(mole_ctx *ctx, uint32_t sys, const mole_operator *obs, uint32_t n);
/* sum over samples of each observable and of its square, walkers summed in
 * a fixed order; *n_samples the sample count. */
mole_status mole_read_observables(mole_ctx *ctx, uint32_t sys, double *sum, double *sum2, uint32_t n,
                                  uint64_t *n_samples);

/* ---- diffusion Monte Carlo -----------------------------------------------
 * DmcRunner (src/dmc/src/dmc.rs:16-203) for every Slater-Jastrow system of a
 * context whose proposal is MOLE_PROPOSAL_DIFFUSE (the MetropolisDiffuse
 * of dmc.rs:28; run.step is its time step).  Each system has its own
 * walkers, weights and reference energy; branching never mixes systems.
 * Streams: the moves continue the walkers' streams from draw 0 of the
 * context epoch at mole_dmc_init; the starting configuration and the
 * branching draws come from streams of their own. */
#define MOLE_BRANCH_SR 0     /* SRBrancher, src/dmc/src/branching.rs:7-44 */
#define MOLE_BRANCH_SIMPLE 1 /* SimpleBranching, branching.rs:46-91 */
/* DmcRunner::new (dmc.rs:45-73): every walker of system k starts from one
 * configuration of N(0,1) coordinates with weight 1; reference_energy[k]
 * is the trial energy.  The walkers' previous configurations are
 * replaced. */
mole_status mole_dmc_init(mole_ctx *ctx, const double *reference_energy, uint32_t brancher);
/* DmcRunner::diffuse (dmc.rs:75-153): num_iterations / block_size blocks
 * of block_size steps; a step moves every electron of every walker once,
 * updates w *= exp(-time_step ((E_old + E_new)/2 - E_ref)), records the
 * ensemble energy sum(w E_old)/sum(w) and branches; after each block from
 * num_eq_blocks on, update_energies (dmc.rs:155-202) mixes E_ref.
 * Unlike the reference's u32 WeightedChoice (which panics once the total
 * integer weight passes 2^32, W > ~65k walkers), SRBrancher totals are
 * 64-bit.  A total integer weight of 0 (the reference panics) returns
 * MOLE_ERR_INVALID_ARGUMENT. */
mole_status mole_dmc_diffuse(mole_ctx *ctx, double time_step, uint32_t num_iterations, uint32_t block_size,
                             uint32_t num_eq_blocks);
/* The (energies, errors) diffuse returns (dmc.rs:147-152) for system sys:
 * one entry per block from num_eq_blocks on, errors sqrt(var_i / (i+1));
 * and the current reference energy.  NULL outputs are skipped. */
mole_status mole_dmc_read(mole_ctx *ctx, uint32_t sys, double *energies, double *errors, size_t cap,
                          size_t *n_blocks, double *reference_energy);
/* The ensemble energy of every step of the last diffuse call (the values
 * dmc.rs:117-121 pushes into each block). */
mole_status mole_dmc_read_steps(mole_ctx *ctx, uint32_t sys, double *step_energy, size_t cap, size_t *n);
/* Walker weights and local energies (DmcRunner::walkers, dmc.rs:27);
 * configurations through mole_get_configs. */
mole_status mole_dmc_read_walkers(mole_ctx *ctx, uint32_t sys, double *weights, double *e_local,
                                  size_t n_walkers);

/* ---- launch configuration ----------------------------------------------
 * No counterpart in the reference (rayon's fan-out, src/vmc/src/vmc.rs:63-76,
 * is the closest).  Selects the two-electron sweep kernel: AUTO picks
 * WAVESPEC where it applies (He, or LCAO with at most 2 nuclei); WAVESPEC
 * gives each 64 walkers a workgroup of four specialised waves (chain, two
 * draw producers, sampler); QUAD spreads each walker over four lanes of a
 * wave (same eligibility);
 * LANE runs one lane per walker with block_threads threads per workgroup
 * (0 keeps the current value).  Every variant produces bit-identical walker
 * trajectories. */
#define MOLE_VARIANT_AUTO 0
#define MOLE_VARIANT_LANE 1
#define MOLE_VARIANT_WAVESPEC 2
#define MOLE_VARIANT_QUAD 3
mole_status mole_set_launch_config(mole_ctx *ctx, uint32_t variant, uint32_t block_threads);
/* Slater-Jastrow: lanes of a wave that share one walker of system sys
 * (power of two, 1..64; 0 = the library's choice).  Results do not depend
 * on it; it only trades per-move parallelism against walkers per wave. */
mole_status mole_set_walker_lanes(mole_ctx *ctx, uint32_t sys, uint32_t lanes);
mole_status mole_get_walker_lanes(mole_ctx *ctx, uint32_t sys, uint32_t *lanes);
/* Slater-Jastrow launch shape: workgroups of one sweep step over all its
 * launches (one launch per LDS footprint class of the tiles) and the
 * largest dynamic LDS bytes per workgroup among them. */
mole_status mole_get_launch_info(mole_ctx *ctx, uint32_t *workgroups, size_t *lds_bytes);
/* Slater-Jastrow tile schedule (no counterpart in the reference).  CLASSES:
 * one launch per LDS footprint class (4 KiB steps) with that class's dynamic
 * LDS, longest tiles first within a class.  PERSISTENT (box proposal, plain
 * kernel; other launches keep CLASSES): one launch of `workgroups` resident
 * workgroups (0: CUs x the occupancy the largest footprint allows) that take
 * every tile of the context, longest first, from a device ticket -- the
 * footprint classes overlap instead of draining one after another, at the
 * largest footprint's LDS.  Results do not depend on the schedule. */
#define MOLE_SCHEDULE_CLASSES 0
#define MOLE_SCHEDULE_PERSISTENT 1
mole_status mole_set_tile_schedule(mole_ctx *ctx, uint32_t schedule, uint32_t workgroups);

/* ---- captured sweeps ---------------------------------------------------
 * No counterpart in the reference (its sweeps are a host loop,
 * src/montecarlo/src/montecarlo.rs:32-33).  Captures one launch of n_sweeps
 * sweeps of every system of a Slater-Jastrow context (sampled with
 * MOLE_SWEEP_SAMPLE), plus the advance of the context's device-resident
 * stream/block state, into a HIP graph.  mole_graph_launch replays it:
 * `replays` replays are exactly `replays` calls of mole_sweep(ctx, n_sweeps,
 * flags).  Parameter changes are seen by replays; a change of lanes per
 * walker makes the graph stale (MOLE_ERR_INVALID_ARGUMENT on launch). */
typedef struct mole_graph mole_graph;
mole_status mole_graph_capture(mole_ctx *ctx, uint32_t n_sweeps, uint32_t flags, mole_graph **out);
mole_status mole_graph_launch(mole_ctx *ctx, mole_graph *graph, uint32_t replays);
void mole_graph_destroy(mole_graph *graph);

/* Waits for all work of the context. */
mole_status mole_synchronize(mole_ctx *ctx);
/* Device time of the kernels of the last mole_sweep/mole_run, measured
 * with HIP events on the context's stream, and the number of sweep-kernel
 * launches it made. */
mole_status mole_last_timing(mole_ctx *ctx, double *ms, uint32_t *launches);

/* ---- learned proposal: network, training data, training ------------------
 * No counterpart in the reference.  The network's inputs for electron i at
 * r: distance to the nearest nucleus, x_j - r for the three nearest other
 * electrons (nearest first; zeros when missing), r.  Outputs: mu = out[0..2],
 * ln sigma = ln 0.25 + 3 tanh(out[3] / 3).  Parameters are FP32 in the
 * order W1[13][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][4]
 * b4[4] (z = b + x W); the kernels run bf16 operands with FP32 accumulation.
 * Training maximises the mean of A |x' - x_i|^2 (expected squared jump
 * distance times acceptance) over (configuration, electron) samples with
 * Adam; the reverse evaluation's inputs are held constant in the gradient. */
#define MOLE_MLP_PARAMS 9476
typedef struct mole_mlp mole_mlp;

/* Training objectives: the expected squared jump distance times acceptance,
 * mean of A |x' - x_i|^2 (ESJD), or the same jump measured against the
 * electron's distance to its nearest nucleus, mean of
 * A |x' - x_i|^2 / (d^2 + 0.01) with d the larger of that distance at x_i
 * and at x' (ESJD_RELATIVE: core and valence electrons weigh alike instead
 * of the valence electrons' large jumps dominating), or the mean of
 * log(1e-4 + A |x' - x_i|^2) (LOG_ESJD: every sample counts by its relative
 * gain, and a rejected move scores as a jump of 0.01 bohr). */
#define MOLE_MLP_OBJ_ESJD 0
#define MOLE_MLP_OBJ_ESJD_RELATIVE 1
#define MOLE_MLP_OBJ_LOG_ESJD 2

typedef struct {
  uint32_t steps;  /* Adam steps */
  uint32_t batch;  /* samples per step, a multiple of 16 */
  float lr, beta1, beta2, eps;
  uint64_t seed;   /* Philox key of the training proposals' normals */
  uint32_t objective; /* MOLE_MLP_OBJ_* */
} mole_train_desc;

/* Xavier-uniform weights from the Philox stream of seed, zero biases. */
mole_status mole_mlp_create(uint64_t seed, mole_mlp **out);
void mole_mlp_destroy(mole_mlp *m);
mole_status mole_mlp_set_weights(mole_mlp *m, const float *w, size_t n);
mole_status mole_mlp_get_weights(mole_mlp *m, float *w, size_t n);
/* the network a MOLE_PROPOSAL_LEARNED context proposes with (not owned;
 * keep it alive while the context sweeps) */
mole_status mole_mlp_attach(mole_ctx *ctx, mole_mlp *m);
/* Runs n_snapshots x sweeps_between unsampled sweeps of ctx and appends the
 * walkers' configurations of system sys after each group to the training
 * set (N <= 16, spin blocks <= 8). */
mole_status mole_mlp_collect(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t n_snapshots,
                             uint32_t sweeps_between);
mole_status mole_mlp_clear_data(mole_mlp *m);
mole_status mole_mlp_data_size(mole_mlp *m, size_t *n_cfg);
/* Adam steps on the training set; sample g (g = step * batch + b) moves
 * electron g mod N of configuration (g / N) mod n_cfg with normals from draws
 * 0..3 of the Philox stream (seed, system 0x6D6C70, lo32(g), hi32(g)).
 * objective (optional, [steps]) receives the batch mean of the training
 * objective d->objective. */
mole_status mole_mlp_train(mole_mlp *m, mole_ctx *ctx, uint32_t sys, const mole_train_desc *d,
                           double *objective);
/* One batch without an update: grad[MOLE_MLP_PARAMS] = sum over the samples
 * first .. first+batch-1 of the derivative of training objective obj_kind
 * (MOLE_MLP_OBJ_*); objective = the sum of the samples' objective values. */
mole_status mole_mlp_gradient(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t batch, uint64_t first,
                              uint64_t seed, uint32_t obj_kind, float *grad, double *objective);
/* Network outputs out[4 e .. 4 e + 3] for electron electron[e] of
 * configuration e of cfg[n][3N] (nuclei nuc_xyz[n_nuc][3]). */
mole_status mole_mlp_eval(mole_mlp *m, const double *cfg, uint32_t N, const double *nuc_xyz, uint32_t n_nuc,
                          const int32_t *electron, uint32_t n, float *out);

#ifdef __cplusplus
}
#endif
#endif /* MOLE_GPU_H */
