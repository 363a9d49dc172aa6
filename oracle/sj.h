/*
 * CPU oracle (test infrastructure only): Slater-Jastrow wavefunctions.
 *
 * The reference ships no determinant wavefunction (README.md:31-33; the
 * removed SpinDeterminantProduct survives only as commented calls at
 * tests/helium_lcao.rs:94-101).  This restates the mathematics the device
 * path implements:
 *
 *   psi = det A_up * det A_dn * exp(J)
 *   A_s[c][k] = phi_k(r of the c-th electron of spin s),
 *   phi_k(r)  = sum_m coef[k][m] chi_m(r - R_center(m))
 *   chi_m     = exp(-zeta r), exp(-zeta r^2) or d_q exp(-zeta r) (kind)
 *   J = sum_{i<j} b_ij r_ij/(1 + c r_ij) - sum_{i,I} a Z_I r_iI/(1 + d r_iI)
 *   b_ij = b_anti (opposite spins) or b_para (same spin)
 *
 * (theory/jastrow.tex:17-30 gives the e-e Pade factor; the e-n factor is
 * the same form.)  Electrons 0..n_up-1 are spin up, the rest spin down.
 *
 * Two runners: orc_sj_run_faithful evaluates psi from scratch twice per
 * move (the reference's accept_move, src/metropolis/src/metrop.rs:73-82)
 * and is the CPU baseline; orc_sj_run_parity performs the
 * Sherman-Morrison algorithm the device runs, in the same operation order
 * and precision, so accept decisions can be compared bit for bit.
 */
#ifndef MOLE_ORACLE_SJ_H
#define MOLE_ORACLE_SJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* b_anti, b_para, c, a, d, and the exponent scale s (live when opt_exp:
 * zeta then holds s zeta0 and d ln psi / d s is formed) */
#define ORC_SJ_NPARAM 6

typedef struct {
  int n_up, n_dn, n_nuc, n_basis, n_orb, hamiltonian;
  const double *nuc;    /* [n_nuc][3] */
  const int *charge;    /* [n_nuc] */
  const double *zeta;   /* [n_basis] */
  const int *center;    /* [n_basis] */
  const double *coef;   /* [n_orb][n_basis] */
  double p[ORC_SJ_NPARAM];
  const int *kind;      /* [n_basis] ORC_BASIS_*; NULL: all STO */
  int opt_exp;          /* 1: p[5] is the exponent scale */
} orc_sj;

/* basis functions about R, d = x - R, r = |d|: exp(-zeta r) (STO),
 * exp(-zeta r^2) (GTO), d_q exp(-zeta r) (p-type STO, q = x, y, z) */
enum { ORC_BASIS_STO = 0, ORC_BASIS_GTO = 1, ORC_BASIS_STO_PX = 2, ORC_BASIS_STO_PY = 3, ORC_BASIS_STO_PZ = 4 };

/* Local operators (O psi)/psi built from terms (LocalOperator::act_on,
 * src/operator/src/traits.rs:263-265, divided by psi as Sampler::sample
 * does): KINETIC -1/2 lap psi/psi (operator.rs:96-121); IONIC_POTENTIAL
 * -sum_{I,i} Z_I/r_iI + ion-ion repulsion (operator.rs:12-64, ions outer,
 * electrons inner); ELECTRONIC_POTENTIAL sum_{i<j} 1/r_ij (operator.rs:
 * 66-94); HARMONIC 1/2 w^2 sum_i |r_i - c|^2 (w = param[0], c = param[1..3];
 * examples/custom_operator.rs:48-61); RADIAL_POWER sum_i |r_i - c|^k
 * (k = param[0] integer, c = param[1..3]); DIPOLE sum_i (r_i - c).u
 * (c = param[0..2], u = param[3..5]).  value = sum_k coef_k term_k. */
enum {
  ORC_TERM_KINETIC = 0,
  ORC_TERM_IONIC_POTENTIAL = 1,
  ORC_TERM_ELECTRONIC_POTENTIAL = 2,
  ORC_TERM_HARMONIC = 3,
  ORC_TERM_RADIAL_POWER = 4,
  ORC_TERM_DIPOLE = 5
};
typedef struct {
  int kind;
  double coef;
  double param[6];
} orc_term;
typedef struct {
  int n_terms;
  orc_term term[4];
} orc_operator;
/* the operator's local value at cfg from the inverses (FP64) */
double orc_sj_operator(const orc_sj *wf, const double *cfg, const double *Bup, const double *Bdn,
                       const orc_operator *op);

enum { ORC_PREC_FP64 = 0, ORC_PREC_MIXED = 1 };

/* Function::value by LU determinants (partial pivoting), FP64. */
double orc_sj_value(const orc_sj *wf, const double *cfg);
/* log|det A_s| and sign, by LU, for tests. */
double orc_sj_logdet(const orc_sj *wf, const double *cfg, int spin, double *sign);
/* Determinant of A_s by the Leibniz expansion in long double (n <= 6). */
long double orc_sj_det_leibniz(const orc_sj *wf, const double *cfg, int spin);
/* Inverse B_s[k][c] = (A_s^-1)[k][c] by Gauss-Jordan with partial
 * pivoting, FP64, exactly as the device refreshes it: a pivot
 * |p| <= n 2^-52 ||A||_1 (or NaN) marks the block singular, its division is
 * skipped (the row is zeroed).  Returns 0, -1 if singular. */
int orc_sj_inverse(const orc_sj *wf, const double *cfg, int spin, double *B);
/* The same, and the determinant from the pivots (row-swap signs times
 * their product) as *mant 2^*ex, the mantissa renormalised by frexp after
 * every factor, as the device carries it; *mant = 0 for a singular block.
 * log|det| = orc_sj_logdet_value(mant, ex). */
int orc_sj_inverse_det(const orc_sj *wf, const double *cfg, int spin, double *B, double *mant, double *ex);
/* log|mant 2^ex| = log|mant| + ex ln2 (ln2 split as in orc_log); -inf for
 * mant = 0. */
double orc_sj_logdet_value(double mant, double ex);
/* mant 2^ex times f, renormalised: the device's per-move update. */
void orc_sj_logdet_mul(double *mant, double *ex, double f);
/* Local energy -1/2 lap(psi)/psi + V and d ln psi / dp of the five
 * Jastrow parameters, FP64, from the inverses B_up, B_dn (as doubles). */
double orc_sj_local_energy(const orc_sj *wf, const double *cfg, const double *Bup,
                           const double *Bdn, double *dlog);
/* The same from scratch (inverses by orc_sj_inverse). */
double orc_sj_local_energy_full(const orc_sj *wf, const double *cfg, double *dlog);

/* drift v = grad_i ln|psi| of electron i (fresh inverse; FP64): the
 * drift_velocity of metrop.rs:153-155 for one row */
void orc_sj_drift(const orc_sj *wf, const double *cfg, int i, double *v);

/* Single-electron move of electron i to xn (fresh FP64 inverse):
 * lR2 = ln (psi(x')/psi(x))^2 and g = grad ln|psi| at x' with respect to
 * electron i. */
void orc_sj_move_terms(const orc_sj *wf, const double *cfg, int i, const double *xn, double *lR2,
                       double *g);

typedef struct {
  uint64_t seed;
  uint32_t sys, epoch;
  double box_side;
  uint32_t steps, block_size;
  int init_from_stream; /* 1: U(-1,1) configurations from draws 0..3N-1 */
  uint64_t pos0;        /* first draw of the moves when !init_from_stream */
  uint64_t move0;       /* global move counter at start (LU refresh phase) */
  int precision;        /* ORC_PREC_* (parity runner only) */
  uint32_t lu_period;   /* refresh both inverses after every lu_period moves */
  int proposal;         /* ORC_SJ_BOX, ORC_SJ_DIFFUSE or ORC_SJ_LEARNED (parity runner only) */
  double tau;           /* drift-diffusion time step */
  const float *mlp_w;   /* learned proposal: network parameters (mlp.h) */
  int mlp_emulate_bf16; /* learned proposal: round operands to bf16 */
  int mlp_drop_q_ratio; /* tests only: a planted mistake (omit q(x|x')/q(x'|x)) */
  const orc_operator *hamiltonian; /* parity runner: the local energy's operator; NULL: built in */
  const orc_operator *obs;         /* parity runner: observables summed over samples */
  int n_obs;
} orc_sj_run_params;

/* Proposals of the parity runner.  BOX: 4 draws per move (3 Range, 1
 * Standard; metrop.rs:60-82).  DIFFUSE (metrop.rs:150-197 with the
 * single-electron Green's function): 6 draws per move -- Box-Muller of draws
 * (0,1) and (2,3) gives z0, z1, z2 (the fourth normal is unused), draw 4 is
 * the accept uniform, draw 5 is unused -- and
 *   x' = (x_i + tau v(x)) + sqrt(tau) z,  v = grad_i ln psi,
 *   reject if the determinant ratio is negative (psi changes sign),
 *   accept iff min(T(x'->x) R^2 / T(x->x'), 1) > u,
 *   T(a->b) = exp(-|b - a - tau v(a)|^2 / (2 tau)) over electron i's row. */
enum { ORC_SJ_BOX = 0, ORC_SJ_DIFFUSE = 1, ORC_SJ_LEARNED = 2 };
/* LEARNED: the draws of DIFFUSE; x' = (x_i + mu) + sigma z with (mu, ln sigma)
 * from the network (mlp.h); accept iff min(R^2 q(x|x')/q(x'|x), 1) > u. */

typedef struct {
  double *cfg;           /* in/out [W][3N] */
  uint64_t *accept_mask; /* [W][blocks*block_size] (N <= 64) */
  double *e_local;       /* [W][samples] */
  double *dlog;          /* [W][samples][5] */
  double sum_e, sum_e2, sum_bm, sum_bm2;
  uint64_t n_samples, n_blocks, n_accepted, n_moves;
  double *obs_sum, *obs_sum2; /* [n_obs]: sums over samples (walkers in order) */
  /* parity runner: the determinants as the device carries them, per walker
   * (mant_up, ex_up, mant_dn, ex_dn) at the end ([W][4], optional); the
   * largest |carried - re-inverted| log|det| at a periodic re-inversion;
   * singular blocks met at re-inversions */
  double *ld;
  double max_ld_drift;
  uint64_t n_singular;
  /* parity runner, optional: the compat "Wavefunction value" samples
   * [W][samples] (orc_sj_compat_psi of the carried determinants) */
  double *psi;
} orc_sj_run_out;

/* "Wavefunction value" of a compat sample as Sampler::sample stores it:
 * (psi psi)/psi (src/vmc/src/operators.rs:19-23, samplers.rs:89-90), psi =
 * sign exp(log|det up| + log|det dn| + J) from the carried determinants ld =
 * (mant_up, ex_up, mant_dn, ex_dn) and J = b_anti dl[0] + b_para dl[1] +
 * a dl[3] from the Jastrow log-derivatives dl (the Pade factor is linear in
 * those three parameters). */
double orc_sj_compat_psi(const orc_sj *wf, const double *ld, const double *dl);

/* Parity runner: walkers [w0, w1), each its own Runner::run
 * (src/montecarlo/src/montecarlo.rs:24-46), Sherman-Morrison ratios and
 * updates, LU refresh every lu_period moves.  Returns 0 or -1. */
int orc_sj_run_parity(const orc_sj *wf, const orc_sj_run_params *rp, uint32_t w0, uint32_t w1,
                      int n_threads, orc_sj_run_out *out);
/* Reference-faithful runner: two full LU evaluations per move and the
 * local energy from fresh inverses after every sampled sweep. */
int orc_sj_run_faithful(const orc_sj *wf, const orc_sj_run_params *rp, uint32_t w0, uint32_t w1,
                        int n_threads, orc_sj_run_out *out);

/* ------------------------------------------------------------------ */
/* Diffusion Monte Carlo (src/dmc/src/dmc.rs, branching.rs)             */
/* ------------------------------------------------------------------ */

/* Philox streams of the DMC steps (key (seed lo32, seed hi32 ^ sys) as for
 * the walkers): the starting configuration is stream (walker 0xFFFFFFFF,
 * epoch ORC_DMC_INIT_EPOCH), draws 0.. in Box-Muller pairs; the branching
 * of step s (counted from mole_dmc_init) is stream (walker s, epoch
 * ORC_DMC_BRANCH_EPOCH).  A uniform integer in [0, n) from a draw d is
 * floor(d n / 2^64). */
#define ORC_DMC_INIT_EPOCH 0x444D4300u
#define ORC_DMC_BRANCH_EPOCH 0x444D4301u
enum { ORC_BRANCH_SR = 0, ORC_BRANCH_SIMPLE = 1 };

/* Sum over walkers in the device's fixed order (blocks of 4096, 256
 * strided partial sums, pairwise trees); the reference folds sequentially,
 * an order that is not pinned and differs only in the last bits. */
double orc_device_sum(const double *x, size_t n);
/* SRBrancher::branch (branching.rs:15-44): integer weights
 * (w W / max w) as u32, WeightedChoice (first index whose cumulative
 * integer weight exceeds floor(d total / 2^64)) for each of the W new
 * walkers, every new weight = mean(w).  src[W] receives the chosen
 * indices, w_out[W] the weights.  Returns -1 for a total integer weight of
 * 0 (the reference panics there). */
int orc_branch_sr(const double *w, uint32_t W, uint64_t seed, uint32_t sys, uint64_t step, uint32_t *src,
                  double *w_out);
/* SimpleBranching::branch (branching.rs:57-91): copies n_i =
 * min(floor(w_i + U_i), 3) (U_i = Standard of draw i), survivors in order,
 * then the births in order; W - M clones of uniformly chosen original
 * walkers (draw W + k), or M - W removals at uniform positions of the
 * shrinking list (draw W + k).  Weights travel with the walkers. */
void orc_branch_simple(const double *w, uint32_t W, uint64_t seed, uint32_t sys, uint64_t step, uint32_t *src,
                       double *w_out);

typedef struct {
  uint64_t seed;
  uint32_t sys, epoch;  /* move stream (walker w, epoch) */
  double tau;           /* MetropolisDiffuse time step (moves) */
  int precision;        /* ORC_PREC_* */
  uint32_t lu_period;
  uint32_t n_walkers;
  int brancher;         /* ORC_BRANCH_* */
  const orc_operator *hamiltonian; /* the local energy's operator; NULL: built in */
} orc_dmc_params;

typedef struct {
  double *cfg;       /* [W][3N] */
  double *weight;    /* [W] */
  double *e_local;   /* [W] local energy at the current configuration */
  double *B;         /* [W][n_up^2 + n_dn^2] inverses (FP64 or FP32 values) */
  double reference_energy;
  uint64_t steps;    /* DMC steps since init (branching stream) */
  uint64_t draw;     /* next draw of the move streams */
  uint64_t moves;    /* moves per walker since init (LU refresh phase) */
} orc_dmc_state;

/* DmcRunner::new (dmc.rs:45-73): every walker starts from the same N(0,1)
 * configuration, weight 1; inverses and local energies of that
 * configuration.  Buffers of st are caller-owned. */
int orc_dmc_init(const orc_sj *wf, const orc_dmc_params *p, double reference_energy, orc_dmc_state *st);
/* DmcRunner::diffuse (dmc.rs:75-153) and update_energies (:155-202):
 * per step, for every walker the cached local energy E_old, one sweep of
 * diffusion moves, E_new, ensemble += w E_old, total += w,
 * w *= exp(-time_step ((E_old + E_new)/2 - E_ref)); the step energy
 * ensemble/total; then the brancher.  energies/errors: one entry per block
 * from block num_eq_blocks on (errors = sqrt(var_i / (i + 1)));
 * step_energy (optional): [num_iterations].  Returns the number of blocks
 * written, or -1. */
int orc_dmc_diffuse(const orc_sj *wf, const orc_dmc_params *p, orc_dmc_state *st, double time_step,
                    uint32_t num_iterations, uint32_t block_size, uint32_t num_eq_blocks, double *energies,
                    double *errors, double *step_energy);

#ifdef __cplusplus
}
#endif
#endif
