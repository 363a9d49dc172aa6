/* Learned MH proposal: CPU restatement of the device MLP (see mlp.h).
 * Test infrastructure only. */
#include "mlp.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define NI ORC_MLP_NIN
#define H ORC_MLP_HID
#define NO ORC_MLP_NOUT

void orc_mlp_features(const orc_sj *wf, const double *cfg, int i, const double *r, double *f) {
  int N = wf->n_up + wf->n_dn;
  double dmin = INFINITY;
  for (int I = 0; I < wf->n_nuc; ++I) {
    double d = orc_norm3(r[0] - wf->nuc[3 * I], r[1] - wf->nuc[3 * I + 1], r[2] - wf->nuc[3 * I + 2]);
    if (d < dmin) dmin = d;
  }
  f[0] = dmin;
  int best[3] = {-1, -1, -1};
  double bd[3] = {INFINITY, INFINITY, INFINITY};
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *x = cfg + 3 * j;
    double d = orc_norm3(x[0] - r[0], x[1] - r[1], x[2] - r[2]);
    for (int k = 0; k < 3; ++k)
      if (d < bd[k]) {
        for (int q = 2; q > k; --q) {
          bd[q] = bd[q - 1];
          best[q] = best[q - 1];
        }
        bd[k] = d;
        best[k] = j;
        break;
      }
  }
  for (int k = 0; k < 3; ++k)
    for (int q = 0; q < 3; ++q) f[1 + 3 * k + q] = best[k] >= 0 ? cfg[3 * best[k] + q] - r[q] : 0.0;
  for (int q = 0; q < 3; ++q) f[10 + q] = r[q];
}

float orc_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
    u = (u | 0x00400000u) & 0xFFFF0000u;
  } else {
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  }
  float y;
  memcpy(&y, &u, 4);
  return y;
}

static double wv(const float *w, int idx, int em) { return em ? (double)orc_bf16(w[idx]) : (double)w[idx]; }
static double av(double x, int em) { return em ? (double)orc_bf16((float)x) : x; }

/* one dense layer z[o] = b[o] + sum_in x[in] W[in][o] */
static void dense(const float *w, int woff, int boff, int nin, int nout, const double *x, double *z) {
  for (int o = 0; o < nout; ++o) z[o] = (double)w[boff + o];
  for (int in = 0; in < nin; ++in)
    for (int o = 0; o < nout; ++o) z[o] += x[in] * (double)w[woff + in * nout + o];
}

void orc_mlp_forward(const float *w, const double *f, int em, double *out, double *act) {
  float wr[ORC_MLP_NPARAM];
  for (int p = 0; p < ORC_MLP_NPARAM; ++p) wr[p] = (float)wv(w, p, em);
  for (int p = ORC_MLP_B1; p < ORC_MLP_B1 + H; ++p) wr[p] = w[p]; /* biases stay FP32 */
  for (int p = ORC_MLP_B2; p < ORC_MLP_B2 + H; ++p) wr[p] = w[p];
  for (int p = ORC_MLP_B3; p < ORC_MLP_B3 + H; ++p) wr[p] = w[p];
  for (int p = ORC_MLP_B4; p < ORC_MLP_B4 + NO; ++p) wr[p] = w[p];
  double x[NI], h1[H], h2[H], h3[H], z[H];
  for (int k = 0; k < NI; ++k) x[k] = av(f[k], em);
  dense(wr, ORC_MLP_W1, ORC_MLP_B1, NI, H, x, z);
  for (int o = 0; o < H; ++o) h1[o] = av(tanh(z[o]), em);
  dense(wr, ORC_MLP_W2, ORC_MLP_B2, H, H, h1, z);
  for (int o = 0; o < H; ++o) h2[o] = av(tanh(z[o]), em);
  dense(wr, ORC_MLP_W3, ORC_MLP_B3, H, H, h2, z);
  for (int o = 0; o < H; ++o) h3[o] = av(tanh(z[o]), em);
  dense(wr, ORC_MLP_W4, ORC_MLP_B4, H, NO, h3, out);
  if (act) {
    memcpy(act, h1, sizeof h1);
    memcpy(act + H, h2, sizeof h2);
    memcpy(act + 2 * H, h3, sizeof h3);
  }
}

/* tanh(out3 / 3) as 1 - 2 / (e^{2y} + 1), the device's formula */
static double sigma_tanh(double out3) {
  double y = fmin(fmax(out3 / ORC_MLP_LS_HALF, -20.0), 20.0);
  return 1.0 - 2.0 / (exp(2.0 * y) + 1.0);
}

double orc_mlp_log_sigma(double out3) { return ORC_MLP_LS0 + ORC_MLP_LS_HALF * sigma_tanh(out3); }

double orc_mlp_log_sigma_slope(double out3) {
  double t = sigma_tanh(out3);
  return 1.0 - t * t;
}

/* backprop of dO/dout through one forward evaluation (FP64), adding into g */
static void backward(const float *w, const double *f, const double *act, const double *dout, int em,
                     double *g) {
  const double *h1 = act, *h2 = act + H, *h3 = act + 2 * H;
  double x[NI], d3[H], d2[H], d1[H];
  for (int k = 0; k < NI; ++k) x[k] = av(f[k], em);
  for (int o = 0; o < NO; ++o) g[ORC_MLP_B4 + o] += dout[o];
  for (int in = 0; in < H; ++in) {
    double s = 0.0;
    for (int o = 0; o < NO; ++o) {
      g[ORC_MLP_W4 + in * NO + o] += h3[in] * dout[o];
      s += wv(w, ORC_MLP_W4 + in * NO + o, em) * dout[o];
    }
    d3[in] = s * (1.0 - h3[in] * h3[in]);
  }
  for (int o = 0; o < H; ++o) g[ORC_MLP_B3 + o] += d3[o];
  for (int in = 0; in < H; ++in) {
    double s = 0.0;
    for (int o = 0; o < H; ++o) {
      g[ORC_MLP_W3 + in * H + o] += h2[in] * d3[o];
      s += wv(w, ORC_MLP_W3 + in * H + o, em) * d3[o];
    }
    d2[in] = s * (1.0 - h2[in] * h2[in]);
  }
  for (int o = 0; o < H; ++o) g[ORC_MLP_B2 + o] += d2[o];
  for (int in = 0; in < H; ++in) {
    double s = 0.0;
    for (int o = 0; o < H; ++o) {
      g[ORC_MLP_W2 + in * H + o] += h1[in] * d2[o];
      s += wv(w, ORC_MLP_W2 + in * H + o, em) * d2[o];
    }
    d1[in] = s * (1.0 - h1[in] * h1[in]);
  }
  for (int o = 0; o < H; ++o) g[ORC_MLP_B1 + o] += d1[o];
  for (int in = 0; in < NI; ++in)
    for (int o = 0; o < H; ++o) g[ORC_MLP_W1 + in * H + o] += x[in] * d1[o];
}

double orc_mlp_sample(const orc_sj *wf, const float *w, const double *cfg, int i, const double *z, int em,
                      const double *frozen_f2, double *grad, double *diag, int objective) {
  int N = wf->n_up + wf->n_dn;
  const double *r = cfg + 3 * i;
  double f1[NI], f2[NI], o1[NO], o2[NO], a1[3 * H], a2[3 * H];
  orc_mlp_features(wf, cfg, i, r, f1);
  orc_mlp_forward(w, f1, em, o1, a1);
  double ls = orc_mlp_log_sigma(o1[3]), sg = exp(ls);
  double xn[3], dx[3];
  for (int q = 0; q < 3; ++q) {
    xn[q] = (r[q] + o1[q]) + sg * z[q];
    dx[q] = o1[q] + sg * z[q];
  }
  if (frozen_f2) {
    memcpy(f2, frozen_f2, sizeof f2);
  } else {
    double cfg2[3 * 64];
    memcpy(cfg2, cfg, sizeof(double) * 3 * N);
    memcpy(cfg2 + 3 * i, xn, sizeof xn);
    orc_mlp_features(wf, cfg2, i, xn, f2);
  }
  orc_mlp_forward(w, f2, em, o2, a2);
  double ls2 = orc_mlp_log_sigma(o2[3]), sg2 = exp(ls2);
  double lR2, gx[3];
  orc_sj_move_terms(wf, cfg, i, xn, &lR2, gx);
  double e[3], e2 = 0.0, z2 = 0.0, d2 = 0.0;
  for (int q = 0; q < 3; ++q) {
    e[q] = (r[q] - xn[q]) - o2[q];
    e2 += e[q] * e[q];
    z2 += z[q] * z[q];
    d2 += dx[q] * dx[q];
  }
  double lq_rev = -3.0 * ls2 - e2 / (2.0 * sg2 * sg2);
  double lq_fwd = -3.0 * ls - 0.5 * z2;
  double logr = lR2 + lq_rev - lq_fwd;
  double A = logr < 0.0 ? exp(logr) : 1.0;
  /* the objective's weight, a constant of the gradient (mlp.h) */
  double wgt = 1.0, O;
  if (objective == 1) {
    double da = (double)(float)f1[0], db = (double)(float)f2[0];
    double dm = da > db ? da : db;
    wgt = 1.0 / (dm * dm + ORC_MLP_REL_D0 * ORC_MLP_REL_D0);
  }
  if (objective == 2) {
    double f = ORC_MLP_LOG_EPS + A * d2;
    wgt = 1.0 / f;
    O = log(f);
  } else {
    O = wgt * (A * d2);
  }
  if (diag) {
    diag[0] = logr;
    diag[1] = A;
    diag[2] = d2;
    diag[3] = ls;
    diag[4] = ls2;
    diag[5] = lR2;
    for (int q = 0; q < 3; ++q) diag[6 + q] = xn[q];
  }
  if (grad) {
    double Aact = (logr < 0.0 ? A : 0.0) * wgt;
    double is2 = 1.0 / (sg2 * sg2);
    double GD[3], dout1[NO], dout2[NO];
    for (int q = 0; q < 3; ++q) GD[q] = 2.0 * (A * wgt) * dx[q] + d2 * Aact * (2.0 * gx[q] + e[q] * is2);
    for (int q = 0; q < 3; ++q) dout1[q] = GD[q];
    double dls = (GD[0] * z[0] + GD[1] * z[1] + GD[2] * z[2]) * sg + d2 * Aact * 3.0;
    dout1[3] = dls * orc_mlp_log_sigma_slope(o1[3]);
    for (int q = 0; q < 3; ++q) dout2[q] = d2 * Aact * (e[q] * is2);
    double dls2 = d2 * Aact * (-3.0 + e2 * is2);
    dout2[3] = dls2 * orc_mlp_log_sigma_slope(o2[3]);
    backward(w, f1, a1, dout1, em, grad);
    backward(w, f2, a2, dout2, em, grad);
  }
  return O;
}

double orc_mlp_log_q_ratio(const orc_sj *wf, const float *w, const double *cfg, int i, const double *xn,
                           int em) {
  int N = wf->n_up + wf->n_dn;
  const double *r = cfg + 3 * i;
  double f1[NI], f2[NI], o1[NO], o2[NO], cfg2[3 * 64];
  orc_mlp_features(wf, cfg, i, r, f1);
  orc_mlp_forward(w, f1, em, o1, NULL);
  memcpy(cfg2, cfg, sizeof(double) * 3 * N);
  memcpy(cfg2 + 3 * i, xn, sizeof(double) * 3);
  orc_mlp_features(wf, cfg2, i, xn, f2);
  orc_mlp_forward(w, f2, em, o2, NULL);
  double ls = orc_mlp_log_sigma(o1[3]), ls2 = orc_mlp_log_sigma(o2[3]);
  double sg = exp(ls), sg2 = exp(ls2);
  double ef = 0.0, er = 0.0;
  for (int q = 0; q < 3; ++q) {
    double a = (xn[q] - r[q]) - o1[q], b = (r[q] - xn[q]) - o2[q];
    ef += a * a;
    er += b * b;
  }
  double lq_fwd = -3.0 * ls - ef / (2.0 * sg * sg);
  double lq_rev = -3.0 * ls2 - er / (2.0 * sg2 * sg2);
  return lq_rev - lq_fwd;
}
