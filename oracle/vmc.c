/*
 * CPU oracle (test infrastructure only): Sampler / Runner / VmcRunner and
 * the optimizer reductions of the reference, restated in C.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Rust f64::min: when one operand is NaN the other is returned
 * (metrop.rs:80 reaches it with 0/0 when psi(x) = 0). */
static double rust_min(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}

typedef struct {
  const orc_wf2 *wf;
  const orc_run_params *rp;
  uint32_t w0, w1;
  orc_run_out *out; /* shared; only per-walker slots are written */
  /* private accumulators */
  double sum_e, sum_e2, sum_bm, sum_bm2;
  double sum_g[ORC_NPARAM2];
  uint64_t n_samples, n_blocks, n_accepted, n_moves;
} run_task;

static void run_walker(run_task *t, uint32_t w) {
  const orc_wf2 *wf = t->wf;
  const orc_run_params *rp = t->rp;
  orc_run_out *out = t->out;
  const int N = 2;
  uint32_t blocks = rp->steps / rp->block_size;
  uint64_t nsamp_w = (uint64_t)(blocks - 1) * rp->block_size;
  uint64_t pos;
  double cfg[6];
  if (rp->init_from_stream) {
    /* Sampler::new: Array2::random_using((nelec, 3), Range(-1, 1)) fills
     * row-major (samplers.rs:46) */
    for (int k = 0; k < 3 * N; ++k)
      cfg[k] = orc_u64_range(
          orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, (uint64_t)k), -1.0,
          1.0);
    pos = 3 * N;
  } else {
    memcpy(cfg, out->cfg + 6 * (size_t)w, sizeof cfg);
    pos = rp->pos0;
  }
  double acceptance = 0.0;
  double half = 0.5 * rp->box_side;
  uint64_t sidx = 0;
  for (uint32_t block = 0; block < blocks; ++block) {
    double bsum = 0.0;
    for (uint32_t s = 0; s < rp->block_size; ++s) {
      /* Sampler::move_state: electrons in order (samplers.rs:106-117) */
      uint64_t mask = 0;
      for (int e = 0; e < N; ++e) {
        /* MetropolisBox::propose_move (metrop.rs:60-71) */
        double prop[6];
        memcpy(prop, cfg, sizeof prop);
        double d[3];
        for (int k = 0; k < 3; ++k)
          d[k] = orc_u64_range(
              orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++), -half,
              half);
        for (int k = 0; k < 3; ++k) prop[3 * e + k] = prop[3 * e + k] + d[k];
        /* MetropolisBox::accept_move (metrop.rs:73-82): two full
         * evaluations, min(psi'^2/psi^2, 1) > U[0,1) */
        double psi_new = orc_wf2_value(wf, prop);
        double psi_old = orc_wf2_value(wf, cfg);
        double a = rust_min((psi_new * psi_new) / (psi_old * psi_old), 1.0);
        double u = orc_u64_standard(
            orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++));
        t->n_moves++;
        if (a > u) {
          memcpy(cfg, prop, sizeof cfg);
          acceptance += 1.0 / (double)N;
          mask |= (uint64_t)1 << e;
          t->n_accepted++;
        }
      }
      if (out->accept_mask)
        out->accept_mask[(size_t)w * blocks * rp->block_size +
                         (size_t)block * rp->block_size + s] = mask;
      /* Runner::run samples after every sweep of blocks > 0
       * (montecarlo.rs:32-38) */
      if (block > 0) {
        double el = orc_wf2_local_energy(wf, cfg);
        size_t si = (size_t)w * nsamp_w + sidx;
        if (out->e_local) out->e_local[si] = el;
        if (out->psi || out->pgrad || rp->all_observables) {
          double psi = orc_wf2_value(wf, cfg);
          /* WavefunctionValue: Scalar(psi^2) / Scalar(psi)
           * (src/vmc/src/operators.rs:19-23, samplers.rs:89-90) */
          if (out->psi) out->psi[si] = (psi * psi) / psi;
          if (out->pgrad || rp->all_observables) {
            /* ParameterGradient: Vector(dpsi) * Scalar(psi), then
             * / Scalar(psi) (operators.rs:9-13); the reference's
             * Vector / Scalar is scalar ./ vector (traits.rs:150). */
            double dp[ORC_NPARAM2];
            orc_wf2_parameter_gradient(wf, cfg, dp);
            for (int k = 0; k < ORC_NPARAM2; ++k) {
              double v = psi * dp[k];
              double G = (rp->o_convention == ORC_O_REFERENCE) ? psi / v : v / psi;
              if (out->pgrad) out->pgrad[si * ORC_NPARAM2 + k] = G;
              t->sum_g[k] += G;
            }
          }
        }
        t->sum_e += el;
        t->sum_e2 += el * el;
        bsum += el;
        t->n_samples++;
        sidx++;
      }
    }
    if (block > 0) {
      double bm = bsum / (double)rp->block_size;
      t->sum_bm += bm;
      t->sum_bm2 += bm * bm;
      t->n_blocks++;
    }
  }
  if (out->cfg) memcpy(out->cfg + 6 * (size_t)w, cfg, sizeof cfg);
  if (out->acceptance) out->acceptance[w] = acceptance;
}

static void *run_task_main(void *arg) {
  run_task *t = (run_task *)arg;
  for (uint32_t w = t->w0; w < t->w1; ++w) run_walker(t, w);
  return NULL;
}

int orc_run_box(const orc_wf2 *wf, const orc_run_params *rp, uint32_t w0,
                uint32_t w1, int n_threads, orc_run_out *out) {
  if (rp->block_size == 0 || rp->steps < 2 * rp->block_size) return -1;
  if (!rp->init_from_stream && !out->cfg) return -1;
  if (n_threads < 1) n_threads = 1;
  uint32_t nw = w1 - w0;
  if ((uint32_t)n_threads > nw) n_threads = nw ? (int)nw : 1;
  run_task *tasks = (run_task *)calloc((size_t)n_threads, sizeof(run_task));
  pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
  for (int i = 0; i < n_threads; ++i) {
    tasks[i].wf = wf;
    tasks[i].rp = rp;
    tasks[i].out = out;
    tasks[i].w0 = w0 + (uint32_t)((uint64_t)nw * i / n_threads);
    tasks[i].w1 = w0 + (uint32_t)((uint64_t)nw * (i + 1) / n_threads);
  }
  if (n_threads == 1) {
    run_task_main(&tasks[0]);
  } else {
    for (int i = 0; i < n_threads; ++i)
      pthread_create(&th[i], NULL, run_task_main, &tasks[i]);
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
  }
  for (int i = 0; i < n_threads; ++i) {
    out->sum_e += tasks[i].sum_e;
    out->sum_e2 += tasks[i].sum_e2;
    out->sum_bm += tasks[i].sum_bm;
    out->sum_bm2 += tasks[i].sum_bm2;
    out->n_samples += tasks[i].n_samples;
    out->n_blocks += tasks[i].n_blocks;
    out->n_accepted += tasks[i].n_accepted;
    out->n_moves += tasks[i].n_moves;
    for (int k = 0; k < ORC_NPARAM2; ++k) out->sum_g[k] += tasks[i].sum_g[k];
  }
  free(tasks);
  free(th);
  return 0;
}

/* ------------------------------------------------------------------ */

double orc_running_mean(const double *x, size_t n) {
  double a = 0.0;
  for (size_t i = 0; i < n; ++i) a = a + (x[i] - a) / (double)(i + 1);
  return a;
}

double orc_blocking_error(const double *x, size_t n, size_t block_size) {
  double mean = orc_running_mean(x, n);
  size_t nblocks = (n + block_size - 1) / block_size; /* chunks() */
  double *bm2 = (double *)malloc(nblocks * sizeof(double));
  for (size_t b = 0; b < nblocks; ++b) {
    size_t lo = b * block_size;
    size_t len = (lo + block_size <= n) ? block_size : n - lo;
    double m = orc_running_mean(x + lo, len);
    bm2[b] = m * m;
  }
  double msq = orc_running_mean(bm2, nblocks);
  free(bm2);
  return sqrt((msq - mean * mean) / (double)(nblocks - 1));
}

void orc_energy_gradient(const double *pgrad, const double *psi,
                         const double *e_local, size_t n, int np,
                         double ebar, double *g) {
  for (int i = 0; i < np; ++i) g[i] = 0.0;
  for (size_t s = 0; s < n; ++s) {
    for (int i = 0; i < np; ++i) {
      double t = pgrad[s * np + i] / psi[s];
      t = t * (e_local[s] - ebar);
      t = 2.0 * t;
      g[i] += t;
    }
  }
  for (int i = 0; i < np; ++i) g[i] = g[i] / (double)n;
}

void orc_sr_matrix(const double *pgrad, const double *psi, size_t n, int np,
                   int sr_mode, double eps, double *S) {
  double *avg = (double *)calloc((size_t)np, sizeof(double));
  double o[16];
  for (int k = 0; k < np * np; ++k) S[k] = 0.0;
  for (size_t s = 0; s < n; ++s) {
    for (int i = 0; i < np; ++i) o[i] = pgrad[s * np + i] / psi[s];
    for (int i = 0; i < np; ++i) {
      avg[i] += o[i];
      for (int j = 0; j < np; ++j) S[i * np + j] += (o[i] * o[j]) / (double)n;
    }
  }
  for (int i = 0; i < np; ++i) avg[i] = avg[i] / (double)n;
  if (sr_mode == ORC_SR_REFERENCE) {
    /* sr_mat -= avg_i avg_j on every element (optimizers.rs:219-224) */
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j)
        for (int k = 0; k < np * np; ++k) S[k] -= avg[i] * avg[j];
    for (int i = 0; i < np; ++i) S[i * np + i] *= 1.0 + 1e-2;
  } else {
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) S[i * np + j] -= avg[i] * avg[j];
    for (int i = 0; i < np; ++i) S[i * np + i] *= 1.0 + eps;
  }
  free(avg);
}

int orc_solve_sym(const double *S, const double *b, int np, double *x) {
  double A[16 * 17];
  int n = np;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) A[i * (n + 1) + j] = S[i * n + j];
    A[i * (n + 1) + n] = b[i];
  }
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * (n + 1) + k]) > fabs(A[piv * (n + 1) + k])) piv = i;
    if (A[piv * (n + 1) + k] == 0.0) return -1;
    if (piv != k)
      for (int j = 0; j <= n; ++j) {
        double tmp = A[k * (n + 1) + j];
        A[k * (n + 1) + j] = A[piv * (n + 1) + j];
        A[piv * (n + 1) + j] = tmp;
      }
    for (int i = k + 1; i < n; ++i) {
      double f = A[i * (n + 1) + k] / A[k * (n + 1) + k];
      for (int j = k; j <= n; ++j) A[i * (n + 1) + j] -= f * A[k * (n + 1) + j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = A[i * (n + 1) + n];
    for (int j = i + 1; j < n; ++j) s -= A[i * (n + 1) + j] * x[j];
    x[i] = s / A[i * (n + 1) + i];
  }
  return 0;
}

int orc_vmc_optimize(orc_wf2 *wf, uint64_t seed, double box_side,
                     uint32_t iters, uint32_t total_samples,
                     uint32_t block_size, uint32_t nworkers,
                     const orc_opt_params *op, int n_threads,
                     double *energies, double *errors, double *acceptance,
                     double *grads, double *params_out) {
  const int np = ORC_NPARAM2;
  uint32_t steps = total_samples / nworkers; /* vmc.rs:50 */
  if (block_size == 0 || steps < 2 * block_size) return -1;
  uint32_t blocks = steps / block_size;
  size_t nsw = (size_t)(blocks - 1) * block_size;
  size_t n = nsw * nworkers;
  double *cfg = (double *)malloc((size_t)nworkers * 6 * sizeof(double));
  double *el = (double *)malloc(n * sizeof(double));
  double *psi = (double *)malloc(n * sizeof(double));
  double *pg = (double *)malloc(n * np * sizeof(double));
  double *acc = (double *)malloc(nworkers * sizeof(double));
  double master[6];
  for (int k = 0; k < 6; ++k)
    master[k] = orc_u64_range(orc_stream_u64(seed, 0, 0xFFFFFFFFu, 0, k),
                              -1.0, 1.0);
  uint64_t pos = 6;
  if (op->persistent) {
    for (uint32_t w = 0; w < nworkers; ++w)
      for (int k = 0; k < 6; ++k)
        cfg[6 * w + k] =
            orc_u64_range(orc_stream_u64(seed, 0, w, 0, k), -1.0, 1.0);
  }
  int rc = 0;
  for (uint32_t it = 0; it < iters; ++it) {
    orc_run_params rp;
    memset(&rp, 0, sizeof rp);
    rp.seed = seed;
    rp.sys = 0;
    rp.box_side = box_side;
    rp.steps = steps;
    rp.block_size = block_size;
    rp.init_from_stream = 0;
    rp.o_convention = op->o_convention;
    if (op->persistent) {
      rp.epoch = 0;
      rp.pos0 = pos;
    } else {
      /* vmc.rs:56: every worker is a clone of the unchanged master */
      for (uint32_t w = 0; w < nworkers; ++w) memcpy(cfg + 6 * w, master, 48);
      rp.epoch = it + 1;
      rp.pos0 = 0;
    }
    orc_run_out out;
    memset(&out, 0, sizeof out);
    out.cfg = cfg;
    out.e_local = el;
    out.psi = psi;
    out.pgrad = pg;
    out.acceptance = acc;
    orc_run_box(wf, &rp, 0, nworkers, n_threads, &out);
    pos += (uint64_t)blocks * block_size * 2 * 4;
    double ebar = orc_running_mean(el, n);
    energies[it] = ebar;
    errors[it] = orc_blocking_error(el, n, block_size);
    double asum = 0.0;
    for (uint32_t w = 0; w < nworkers; ++w) asum += acc[w];
    acceptance[it] = asum / (double)total_samples;
    double g[ORC_NPARAM2], dp[ORC_NPARAM2];
    orc_energy_gradient(pg, psi, el, n, np, ebar, g);
    if (op->optimizer == ORC_OPT_SD) {
      for (int i = 0; i < np; ++i) dp[i] = -(op->step * g[i]);
    } else {
      double S[ORC_NPARAM2 * ORC_NPARAM2], rhs[ORC_NPARAM2], x[ORC_NPARAM2];
      orc_sr_matrix(pg, psi, n, np, op->sr_mode, op->sr_eps, S);
      for (int i = 0; i < np; ++i) rhs[i] = -0.5 * g[i];
      if (orc_solve_sym(S, rhs, np, x) != 0) {
        rc = -2;
        break;
      }
      for (int i = 0; i < np; ++i) dp[i] = op->step * x[i];
    }
    for (int i = 0; i < np; ++i) {
      grads[(size_t)it * np + i] = g[i];
      wf->p[i] = wf->p[i] + dp[i];
      params_out[(size_t)it * np + i] = wf->p[i];
    }
  }
  free(cfg);
  free(el);
  free(psi);
  free(pg);
  free(acc);
  return rc;
}

/* ------------------------------------------------------------------ */

double orc_tau_corr(const double *x, size_t n) {
  double mean = 0.0;
  for (size_t i = 0; i < n; ++i) mean += x[i];
  mean /= (double)n;
  double var = 0.0;
  for (size_t i = 0; i < n; ++i) var += (mean - x[i]) * (mean - x[i]);
  var /= (double)(n - 1);
  const size_t MAX_STEPS = 200;
  size_t max_i = (MAX_STEPS < n - 1) ? MAX_STEPS : n - 1;
  double tau = 1.0, f = 1.0;
  for (size_t i = 1; i <= max_i; ++i) {
    double corr = 0.0;
    for (size_t t = 0; t + i < n; ++t) corr += (x[t] - mean) * (x[t + i] - mean);
    corr /= var * (double)(n - i);
    if (corr < 0.0) f = 0.0;
    tau += 2.0 * corr * f;
  }
  return tau > 1.0 ? tau : 1.0;
}

/* ------------------------------------------------------------------ */
/* MomentumDescent, NesterovMomentum, OnlineLbfgs (optimizers.rs:32-179) */
/* ------------------------------------------------------------------ */

int orc_stateful_opt_init(orc_stateful_opt *o, int kind, int np, double step, double mu, int history) {
  if (np < 1 || np > 16) return -1;
  if (kind == ORC_OPT_LBFGS && (history < 1 || history > ORC_MAX_HISTORY)) return -1;
  memset(o, 0, sizeof(*o));
  o->kind = kind;
  o->np = np;
  o->step = step;
  o->mu = mu;
  o->history = history;
  /* OnlineLbfgs::new: grad_prev and pars_prev are filled with EPS = 1e-5 */
  for (int i = 0; i < np; ++i) {
    o->grad_prev[i] = 1e-5;
    o->pars_prev[i] = 1e-5;
  }
  return 0;
}

static double vdot(const double *a, const double *b, int n) {
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += a[i] * b[i];
  return t;
}

/* update_curvature_pairs (optimizers.rs:146-158) */
static void lbfgs_pairs(orc_stateful_opt *o, const double *pars, const double *g) {
  const int n = o->np;
  if (o->len >= o->history) { /* pop_front */
    for (int k = 1; k < o->len; ++k) {
      memcpy(o->s[k - 1], o->s[k], sizeof(o->s[0]));
      memcpy(o->y[k - 1], o->y[k], sizeof(o->y[0]));
    }
    o->len -= 1;
  }
  for (int i = 0; i < n; ++i) {
    o->s[o->len][i] = pars[i] - o->pars_prev[i];
    o->y[o->len][i] = g[i] - o->grad_prev[i];
  }
  o->len += 1;
}

/* initial_direction (optimizers.rs:121-144) */
static void lbfgs_direction(const orc_stateful_opt *o, const double *g, double *p) {
  const int n = o->np;
  double alphas[ORC_MAX_HISTORY];
  int na = 0;
  for (int i = 0; i < n; ++i) p[i] = -g[i];
  for (int k = o->len - 1; k >= 0; --k) { /* s.iter().rev().zip(y.iter().rev()) */
    const double alpha = vdot(o->s[k], p, n) / vdot(o->s[k], o->y[k], n);
    for (int i = 0; i < n; ++i) p[i] -= alpha * o->y[k][i];
    alphas[na++] = alpha;
  }
  double scale;
  if (o->iter == 0) {
    scale = 1e-10;
  } else {
    double tot = 0.0;
    for (int k = 0; k < o->len; ++k) tot = tot + vdot(o->s[k], o->y[k], n) / vdot(o->y[k], o->y[k], n);
    const size_t d = o->iter < (size_t)o->history ? o->iter : (size_t)o->history;
    scale = tot / (double)d;
  }
  for (int i = 0; i < n; ++i) p[i] *= scale;
  /* izip!(alphas.iter().rev(), s.iter(), y.iter()) */
  for (int k = 0; k < o->len; ++k) {
    const double alpha = alphas[na - 1 - k];
    const double c = alpha - vdot(o->y[k], p, n) / vdot(o->y[k], o->s[k], n);
    for (int i = 0; i < n; ++i) p[i] = p[i] + c * o->s[k][i];
  }
}

void orc_stateful_opt_update(orc_stateful_opt *o, const double *pars, const double *g, double *dp) {
  const int n = o->np;
  if (o->kind == ORC_OPT_MOMENTUM) {
    /* self.momentum -= &(self.step_size * energy_grad);
       Ok(self.momentum_parameter * &self.momentum) */
    for (int i = 0; i < n; ++i) {
      o->momentum[i] -= o->step * g[i];
      dp[i] = o->mu * o->momentum[i];
    }
  } else if (o->kind == ORC_OPT_NESTEROV) {
    memcpy(o->momentum_prev, o->momentum, sizeof(o->momentum));
    for (int i = 0; i < n; ++i) o->momentum[i] = o->mu * o->momentum[i] + o->step * g[i];
    for (int i = 0; i < n; ++i) dp[i] = -(o->mu * o->momentum_prev[i] + (1.0 + o->mu) * o->momentum[i]);
  } else {
    double p[16];
    lbfgs_pairs(o, pars, g);
    lbfgs_direction(o, g, p);
    for (int i = 0; i < n; ++i) dp[i] = -o->step * p[i];
    lbfgs_pairs(o, pars, g);
    memcpy(o->grad_prev, g, sizeof(double) * n);
    memcpy(o->pars_prev, pars, sizeof(double) * n);
  }
  o->iter += 1;
}
