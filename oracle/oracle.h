/*
 * CPU oracle for the mole walker hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's algorithms (Jvanrhijn/mole, Rust).  It is linked or loaded
 * only by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg.  The product (mole_amd/, libmole_amd.so)
 * never includes, links or calls anything in this directory.
 *
 * Every entry point cites the reference file:line whose behaviour it
 * restates.  Where the reference has no code (it ships no wavefunctions,
 * no Slater determinants, no Philox), the function states the mathematics
 * it implements and the checks that pin it live in tests/.
 *
 * Floating point: compiled with -ffp-contract=off and no fast-math, so
 * every a*b+c rounds twice unless fma() is written explicitly.  Rust never
 * contracts, so mul+add here reproduces the reference's own rounding
 * wherever the reference spells the arithmetic out.
 */
#ifndef MOLE_ORACLE_H
#define MOLE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10 and the per-walker u64 draw stream                    */
/* ------------------------------------------------------------------ */

/* Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11),
 * Philox4x32 with 10 rounds.  Replaces the reference's StdRng
 * (src/metropolis/src/metrop.rs:5,43) behind the pluggable `R: Rng`
 * bound of src/metropolis/src/traits.rs:10-13. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

/* Draw number i of the u64 stream of (walker, epoch) in system sys:
 *   key = (lo32(seed), hi32(seed) ^ sys)
 *   ctr = (lo32(i>>1), hi32(i>>1), walker, epoch)
 *   u64 = word[2j] | word[2j+1] << 32, j = i & 1
 * which is how a 32-bit block RNG hands out next_u64 (low word first). */
uint64_t orc_stream_u64(uint64_t seed, uint32_t sys, uint32_t walker,
                        uint32_t epoch, uint64_t i);

/* rand 0.5 `Standard` for f64: (x >> 11) * 2^-53, in [0, 1).
 * Used by metrop.rs:81 (`self.rng.gen::<f64>()`). */
double orc_u64_standard(uint64_t x);

/* rand 0.5 `Range::new(lo, hi)` for f64: v in [1,2) from the top 52 bits,
 * then v*scale + offset with scale = hi-lo, offset = lo-scale (two
 * roundings).  Used by metrop.rs:64-68 and samplers.rs:46. */
double orc_u64_range(uint64_t x, double lo, double hi);

/* ------------------------------------------------------------------ */
/* Elementary functions shared bit-for-bit with the device             */
/* ------------------------------------------------------------------ */

/* exp(x): Cody-Waite reduction x = k ln2 + r with the fdlibm split of
 * ln2, degree-13 Taylor polynomial of e^r by Horner with fma, scaled by
 * 2^k in two exact steps.  Stands in for Rust's f64::exp
 * (examples/helium_atom_singlet.rs:69). */
double orc_exp(double x);

/* |v|_2 as ndarray-linalg's norm_l2 spells it: sqrt((x*x + y*y) + z*z). */
double orc_norm3(double x, double y, double z);

/* ln x (x > 0, normal): atanh series of (m-1)/(m+1) after the exponent is
 * split off; within a few ulp of the correctly rounded log. */
double orc_log(double x);

/* sin(pi x), cos(pi x) for 0 <= x < 2 by quadrant reduction and Taylor
 * polynomials on [-pi/4, pi/4]. */
void orc_sincospi(double x, double *sn, double *cs);

/* Two N(0,1) draws from two u64 draws (Box-Muller with u1 = 1 - Standard,
 * so u1 > 0).  Stands in for rand 0.5's ziggurat Normal (metrop.rs:160),
 * whose stream cannot be reproduced here. */
void orc_box_muller(uint64_t a, uint64_t b, double *z0, double *z1);

/* ------------------------------------------------------------------ */
/* Two-electron wavefunctions (configs: He singlet, H2 LCAO)           */
/* ------------------------------------------------------------------ */

enum {
  ORC_WF_HELIUM_PADE = 1, /* psi = exp(-zeta (r1 + r2) + b r12/(1 + c r12)) */
  ORC_WF_LCAO2_PADE = 2   /* psi = phi(r1) phi(r2) exp(b r12/(1 + c r12)),
                             phi(r) = sum_I exp(-zeta |r - R_I|)          */
};
enum { ORC_H_ELECTRONIC = 0, ORC_H_IONIC = 1 };
enum { ORC_O_LOGDERIV = 0, ORC_O_REFERENCE = 1 };
enum { ORC_SR_TEXTBOOK = 0, ORC_SR_REFERENCE = 1 };
enum { ORC_OPT_SD = 0, ORC_OPT_SR = 1 };

#define ORC_MAX_NUC 8
#define ORC_NPARAM2 3 /* (zeta, b, c) */

typedef struct {
  int kind;        /* ORC_WF_* */
  int hamiltonian; /* ORC_H_*  */
  int n_nuc;
  double nuc[ORC_MAX_NUC][3];
  int charge[ORC_MAX_NUC];
  double p[ORC_NPARAM2]; /* zeta, b, c */
} orc_wf2;

/* Function::value (src/wavefunction_traits/src/lib.rs:7-11). cfg is the
 * (N=2, 3) row-major configuration. */
double orc_wf2_value(const orc_wf2 *wf, const double cfg[6]);
/* Differentiate::gradient / laplacian (lib.rs:14-19): gradient of psi (not
 * of ln psi) and the laplacian of psi, as the reference's user code returns
 * them (examples/helium_atom_singlet.rs:76-98). */
void orc_wf2_gradient(const orc_wf2 *wf, const double cfg[6], double g[6]);
double orc_wf2_laplacian(const orc_wf2 *wf, const double cfg[6]);
/* Optimize::parameter_gradient (src/optimize/src/traits.rs:9):
 * d psi / d p for p = (zeta, b, c). */
void orc_wf2_parameter_gradient(const orc_wf2 *wf, const double cfg[6],
                                double dpsi[ORC_NPARAM2]);
/* IonicPotential::value (src/operator/src/operator.rs:25-36) including the
 * nuclear repulsion of IonicPotential::new (operator.rs:40-55). */
double orc_ionic_potential(const orc_wf2 *wf, const double *cfg, int nelec);
/* ElectronicPotential::value (operator.rs:80-90). */
double orc_electronic_potential(const double *cfg, int nelec);
/* Sampler::sample of the "Energy" observable: ElectronicHamiltonian::act_on
 * (operator.rs:177-183, or IonicHamiltonian operator.rs:142-149) divided by
 * Scalar(psi) (src/montecarlo/src/samplers.rs:89-90). */
double orc_wf2_local_energy(const orc_wf2 *wf, const double cfg[6]);

/* ------------------------------------------------------------------ */
/* Sampler / Runner restatement (box proposal)                         */
/* ------------------------------------------------------------------ */

typedef struct {
  uint64_t seed;
  uint32_t sys;        /* system index mixed into the Philox key */
  uint32_t epoch;      /* Philox counter word 3 */
  double box_side;     /* MetropolisBox::box_side (metrop.rs:26,66) */
  uint32_t steps;      /* Runner::run steps (montecarlo.rs:24-46) */
  uint32_t block_size; /* Runner::run block_size */
  int init_from_stream;/* 1: Sampler::new draws cfg (samplers.rs:46) from
                          draws 0..3N-1 and moves start at draw 3N;
                          0: cfg given, moves start at draw pos0 */
  uint64_t pos0;
  int o_convention;    /* ORC_O_* for the "Parameter gradient" samples */
  int all_observables; /* 1: also evaluate "Wavefunction value" and
                          "Parameter gradient" every sample, even when they
                          are not stored (the VMC-optimisation cost) */
} orc_run_params;

/* Optional per-walker outputs of a run; any pointer may be NULL.
 * Sample arrays are walker-major: [walker][sample] (the order in which
 * VmcRunner::concatenate_worker_data appends worker data, vmc.rs:108-130).
 * samples per walker = (blocks - 1) * block_size. */
typedef struct {
  double *cfg;            /* in/out [W][6]; in when !init_from_stream */
  uint64_t *accept_mask;  /* [W][blocks*block_size] bit e = move of e accepted */
  double *e_local;        /* [W][samples] "Energy" */
  double *psi;            /* [W][samples] "Wavefunction value" (= psi)  */
  double *pgrad;          /* [W][samples][3] "Parameter gradient" as stored */
  double *acceptance;     /* [W] Sampler::acceptance (samplers.rs:113) */
  /* Accumulated statistics over all walkers (for rate runs without
   * per-sample storage): sums of E and E^2, of block means and squares. */
  double sum_e, sum_e2, sum_bm, sum_bm2;
  double sum_g[ORC_NPARAM2]; /* sums of the stored-form parameter gradient */
  uint64_t n_samples, n_blocks, n_accepted, n_moves;
} orc_run_out;

/* Runs walkers [w0, w1) of a batch, each as its own Sampler + Runner::run
 * (samplers.rs:40-117, montecarlo.rs:24-46, metrop.rs:60-96), with the
 * reference's cost structure: two full psi evaluations per move and the
 * generic act_on/psi local energy after every sampled sweep.  Walker w
 * uses Philox stream (w, epoch).  n_threads > 1 spreads walkers over
 * POSIX threads like rayon's fan-out (src/vmc/src/vmc.rs:63-76).
 * Returns 0, or -1 when steps < 2*block_size (the assert at
 * montecarlo.rs:29). */
int orc_run_box(const orc_wf2 *wf, const orc_run_params *rp, uint32_t w0,
                uint32_t w1, int n_threads, orc_run_out *out);

/* ------------------------------------------------------------------ */
/* Reductions: running mean, blocking error, energy gradient, SR, SD   */
/* ------------------------------------------------------------------ */

/* VmcRunner::mean (vmc.rs:174-179): a += (b - a)/(n + 1). */
double orc_running_mean(const double *x, size_t n);
/* Blocking error of process_monte_carlo_results (vmc.rs:150-170):
 * sqrt((mean(B^2) - mean^2) / (nblocks - 1)), chunks of block_size,
 * with mean = running mean of all samples. */
double orc_blocking_error(const double *x, size_t n, size_t block_size);
/* compute_energy_gradient (src/optimize/src/util.rs:6-46): rows
 * 2*(G/psi)*(E - Ebar), then the column mean.  G are the stored
 * "Parameter gradient" samples, psi the "Wavefunction value" samples. */
void orc_energy_gradient(const double *pgrad, const double *psi,
                         const double *e_local, size_t n, int np,
                         double ebar, double *g);
/* StochasticReconfiguration::construct_sr_matrix (optimizers.rs:191-233)
 * in ORC_SR_REFERENCE mode (Q2: scalar subtraction on every element,
 * diag *= 1 + 1e-2) or ORC_SR_TEXTBOOK mode (S_ij -= O_i O_j,
 * diag *= 1 + eps). */
void orc_sr_matrix(const double *pgrad, const double *psi, size_t n, int np,
                   int sr_mode, double eps, double *S);
/* Solve S x = b for symmetric S (stand-in for ndarray-linalg solveh_into,
 * optimizers.rs:251): Gaussian elimination with partial pivoting.
 * Returns 0, or -1 for a singular matrix (LinalgError). */
int orc_solve_sym(const double *S, const double *b, int np, double *x);

/* VmcRunner::run_optimization (vmc.rs:43-106) restated for the
 * two-electron wavefunctions.  Every iteration restarts all nworkers
 * walkers from the master configuration (vmc.rs:56) and runs Runner::run
 * with steps = total_samples / nworkers (vmc.rs:50).  The worker RNGs of
 * iteration it are Philox streams (worker, epoch = it + 1); the master
 * configuration comes from stream (0xFFFFFFFF, 0).  When persistent != 0
 * the walkers instead keep their configurations and streams between
 * iterations (the device-native schedule) and start from their own
 * stream-drawn configurations.
 * Outputs per iteration: energy, blocking error, acceptance
 * (= accepted/(N*total sweeps), vmc.rs:97), energy gradient g[np] and the
 * parameters after the update. */
typedef struct {
  int optimizer;    /* ORC_OPT_SD or ORC_OPT_SR */
  double step;      /* SteepestDescent::step_size / SR step_size */
  int sr_mode;      /* ORC_SR_* */
  double sr_eps;    /* textbook diagonal shift */
  int o_convention; /* ORC_O_* */
  int persistent;
} orc_opt_params;

int orc_vmc_optimize(orc_wf2 *wf, uint64_t seed, double box_side,
                     uint32_t iters, uint32_t total_samples,
                     uint32_t block_size, uint32_t nworkers,
                     const orc_opt_params *op, int n_threads,
                     double *energies, double *errors, double *acceptance,
                     double *grads, double *params_out);

/* Stateful optimizers of src/optimize/src/optimizers.rs, restated on a
 * given energy gradient g (as compute_energy_gradient returns it):
 * MomentumDescent (:32-61), NesterovMomentum (:63-95), OnlineLbfgs
 * (:97-179, with its double update_curvature_pairs at :163,:170).  The
 * LBFGS pair lists are VecDeques: oldest first, pop_front when full. */
#define ORC_MAX_HISTORY 64
typedef struct {
  int kind; /* ORC_OPT_MOMENTUM / ORC_OPT_NESTEROV / ORC_OPT_LBFGS */
  int np;
  double step, mu;
  int history;
  size_t iter;
  double momentum[16], momentum_prev[16];
  double grad_prev[16], pars_prev[16];
  int len;
  double s[ORC_MAX_HISTORY][16], y[ORC_MAX_HISTORY][16];
} orc_stateful_opt;
#define ORC_OPT_MOMENTUM 2
#define ORC_OPT_NESTEROV 3
#define ORC_OPT_LBFGS 4
/* ::new of each (optimizers.rs:40-46, :71-78, :108-119).  Returns -1 for
 * np outside 1..16 or history outside 1..ORC_MAX_HISTORY. */
int orc_stateful_opt_init(orc_stateful_opt *o, int kind, int np, double step, double mu, int history);
/* compute_parameter_update after the gradient: writes dp[np]. */
void orc_stateful_opt_update(orc_stateful_opt *o, const double *pars, const double *g, double *dp);

/* ------------------------------------------------------------------ */
/* Statistics of scripts/statfor.rs                                    */
/* ------------------------------------------------------------------ */

/* correlation() (scripts/statfor.rs:30-54): tau = max(1, 1 + 2 sum_i rho_i)
 * over i <= min(200, n-1), terms zeroed from the first negative rho on;
 * rho_i = sum (x_t - m)(x_{t+i} - m) / (var (n - i)). var is ddof=1
 * (statfor.rs:88).  Returns tau. */
double orc_tau_corr(const double *x, size_t n);

#ifdef __cplusplus
}
#endif
#endif
