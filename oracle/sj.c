/*
 * CPU oracle (test infrastructure only): Slater-Jastrow wavefunction, its
 * local energy, and the two Metropolis runners described in sj.h.
 */
#include "sj.h"
#include "mlp.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXE 64
#define MAXN 32
#define MAXNUC 64
#define MAXB 512

static int spin_of(const orc_sj *wf, int i) { return i < wf->n_up ? 0 : 1; }
static int col_of(const orc_sj *wf, int i) { return i < wf->n_up ? i : i - wf->n_up; }
static int nspin(const orc_sj *wf, int s) { return s == 0 ? wf->n_up : wf->n_dn; }
static int electron_of(const orc_sj *wf, int s, int c) { return s == 0 ? c : wf->n_up + c; }

static double rust_min1(double a) { return (a < 1.0) ? a : 1.0; }

static int kind_of(const orc_sj *wf, int m) { return wf->kind ? wf->kind[m] : ORC_BASIS_STO; }

/* Value v, gradient g and laplacian l of a basis function of the given kind
 * at d = x - R, r = |d| (the device's basis_vgl, same operations):
 *   STO    e = exp(-z r), g = (-z e / r) d, l = e (z^2 - 2 z/r)
 *   GTO    v = exp(-z r^2), g = (-2 z v) d, l = v (4 z^2 r^2 - 6 z)
 *   STO_P  v = d_q e, g_k = ((-z e / r) d_q) d_k + (k == q ? e : 0),
 *          l = v (z^2 - 4 z/r)  (lap(x f(r)) = x (f'' + 4 f'/r))
 * with 1/r formed once and multiplied. */
static void basis_vgl(int kind, double z, const double d[3], double r, double *v, double g[3], double *l) {
  if (kind == ORC_BASIS_GTO) {
    const double r2 = r * r;
    *v = orc_exp(-z * r2);
    const double gf = -2.0 * z * *v;
    for (int q = 0; q < 3; ++q) g[q] = gf * d[q];
    *l = *v * (4.0 * z * z * r2 - 6.0 * z);
    return;
  }
  const double e = orc_exp(-z * r);
  const double ir = 1.0 / r; /* one division, the device's */
  const double gf = (-z * e) * ir;
  if (kind >= ORC_BASIS_STO_PX) {
    const int q = kind - ORC_BASIS_STO_PX;
    const double gq = gf * d[q];
    *v = d[q] * e;
    for (int k = 0; k < 3; ++k) g[k] = gq * d[k] + (k == q ? e : 0.0);
    *l = *v * (z * z - (4.0 * z) * ir);
  } else {
    *v = e;
    for (int k = 0; k < 3; ++k) g[k] = gf * d[k];
    *l = e * (z * z - (2.0 * z) * ir);
  }
}

/* basis values chi[m] at r (distances to all nuclei first) */
static void basis_at(const orc_sj *wf, const double *r, double *chi) {
  double rn[MAXNUC];
  for (int I = 0; I < wf->n_nuc; ++I)
    rn[I] = orc_norm3(r[0] - wf->nuc[3 * I], r[1] - wf->nuc[3 * I + 1], r[2] - wf->nuc[3 * I + 2]);
  for (int m = 0; m < wf->n_basis; ++m) {
    const int I = wf->center[m];
    const double rr = rn[I];
    const int kd = kind_of(wf, m);
    if (kd == ORC_BASIS_GTO) {
      chi[m] = orc_exp(-wf->zeta[m] * (rr * rr));
    } else if (kd >= ORC_BASIS_STO_PX) {
      const int q = kd - ORC_BASIS_STO_PX;
      chi[m] = (r[q] - wf->nuc[3 * I + q]) * orc_exp(-wf->zeta[m] * rr);
    } else {
      chi[m] = orc_exp(-wf->zeta[m] * rr);
    }
  }
}

static double orbital(const orc_sj *wf, int k, const double *chi) {
  const double *c = wf->coef + (size_t)k * wf->n_basis;
  double phi = 0.0;
  for (int m = 0; m < wf->n_basis; ++m) phi = fma(c[m], chi[m], phi);
  return phi;
}

/* Slater matrix A_s[c][k] (FP64) */
static void slater_matrix(const orc_sj *wf, const double *cfg, int s, double *A) {
  int n = nspin(wf, s);
  double chi[MAXB];
  for (int c = 0; c < n; ++c) {
    basis_at(wf, cfg + 3 * electron_of(wf, s, c), chi);
    for (int k = 0; k < n; ++k) A[c * n + k] = orbital(wf, k, chi);
  }
}

double orc_sj_logdet(const orc_sj *wf, const double *cfg, int spin, double *sign) {
  int n = nspin(wf, spin);
  double A[MAXN * MAXN];
  slater_matrix(wf, cfg, spin, A);
  double sg = 1.0, ld = 0.0;
  for (int p = 0; p < n; ++p) {
    int piv = p;
    for (int r = p + 1; r < n; ++r)
      if (fabs(A[r * n + p]) > fabs(A[piv * n + p])) piv = r;
    if (A[piv * n + p] == 0.0) {
      *sign = 0.0;
      return -INFINITY;
    }
    if (piv != p) {
      for (int j = 0; j < n; ++j) {
        double t = A[p * n + j];
        A[p * n + j] = A[piv * n + j];
        A[piv * n + j] = t;
      }
      sg = -sg;
    }
    double pv = A[p * n + p];
    if (pv < 0) sg = -sg;
    ld += log(fabs(pv));
    for (int r = p + 1; r < n; ++r) {
      double f = A[r * n + p] / pv;
      for (int j = p; j < n; ++j) A[r * n + j] -= f * A[p * n + j];
    }
  }
  *sign = sg;
  return ld;
}

static double lu_det(const orc_sj *wf, const double *cfg, int s) {
  double sg;
  double ld = orc_sj_logdet(wf, cfg, s, &sg);
  return sg == 0.0 ? 0.0 : sg * exp(ld);
}

long double orc_sj_det_leibniz(const orc_sj *wf, const double *cfg, int spin) {
  int n = nspin(wf, spin);
  double A[MAXN * MAXN];
  slater_matrix(wf, cfg, spin, A);
  if (n == 0) return 1.0L;
  /* iterate permutations with Heap's algorithm */
  int perm[8], cnt[8];
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    cnt[i] = 0;
  }
  int sgn = 1;
  long double total = 0.0L;
  {
    long double t = sgn;
    for (int r = 0; r < n; ++r) t *= A[r * n + perm[r]];
    total += t;
  }
  int i = 0;
  while (i < n) {
    if (cnt[i] < i) {
      int a = (i % 2 == 0) ? 0 : cnt[i];
      int tmp = perm[a];
      perm[a] = perm[i];
      perm[i] = tmp;
      sgn = -sgn;
      long double t = sgn;
      for (int r = 0; r < n; ++r) t *= A[r * n + perm[r]];
      total += t;
      cnt[i]++;
      i = 0;
    } else {
      cnt[i] = 0;
      i++;
    }
  }
  return total;
}

double orc_sj_logdet_value(double mant, double ex) {
  if (mant == 0.0) return -INFINITY;
  /* ln 2 split as in orc_log (fdlibm) */
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  return fma(ex, ln2_hi, fma(ex, ln2_lo, orc_log(fabs(mant))));
}

double orc_sj_compat_psi(const orc_sj *wf, const double *ld, const double *dl) {
  const double J = (wf->p[0] * dl[0] + wf->p[1] * dl[1]) + wf->p[3] * dl[3];
  const double lg = (orc_sj_logdet_value(ld[0], ld[1]) + orc_sj_logdet_value(ld[2], ld[3])) + J;
  const double sg = ((ld[0] < 0.0) != (ld[2] < 0.0)) ? -1.0 : 1.0;
  const double psi = sg * orc_exp(lg);
  return (psi * psi) / psi;
}

void orc_sj_logdet_mul(double *mant, double *ex, double f) {
  int e;
  *mant = frexp(*mant * f, &e);
  *ex = *ex + (double)e;
}

int orc_sj_inverse_det(const orc_sj *wf, const double *cfg, int spin, double *B, double *mant, double *ex) {
  int n = nspin(wf, spin);
  double M[MAXN * MAXN], I[MAXN * MAXN];
  slater_matrix(wf, cfg, spin, M);
  /* ||A||_1: column sums in row order, their maximum */
  double anorm = 0.0;
  for (int k = 0; k < n; ++k) {
    double cs = 0.0;
    for (int c = 0; c < n; ++c) cs += fabs(M[c * n + k]);
    anorm = cs > anorm ? cs : anorm;
  }
  const double thresh = anorm * ((double)n * 0x1p-52);
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < n; ++j) I[r * n + j] = (r == j) ? 1.0 : 0.0;
  int rc = 0;
  double dm = 1.0, de = 0.0;
  for (int p = 0; p < n; ++p) {
    int piv = p;
    for (int r = p + 1; r < n; ++r)
      if (fabs(M[r * n + p]) > fabs(M[piv * n + p])) piv = r;
    const double pv0 = M[piv * n + p];
    const int bad = !(fabs(pv0) > thresh);
    if (bad) {
      rc = -1;
    } else {
      if (piv != p) dm = -dm;
      orc_sj_logdet_mul(&dm, &de, pv0);
    }
    if (piv != p)
      for (int j = 0; j < n; ++j) {
        double t = M[p * n + j];
        M[p * n + j] = M[piv * n + j];
        M[piv * n + j] = t;
        t = I[p * n + j];
        I[p * n + j] = I[piv * n + j];
        I[piv * n + j] = t;
      }
    double ip = bad ? 0.0 : 1.0 / M[p * n + p];
    for (int j = 0; j < n; ++j) {
      M[p * n + j] = M[p * n + j] * ip;
      I[p * n + j] = I[p * n + j] * ip;
    }
    for (int r = 0; r < n; ++r) {
      if (r == p) continue;
      double f = M[r * n + p];
      for (int j = 0; j < n; ++j) {
        M[r * n + j] = fma(-f, M[p * n + j], M[r * n + j]);
        I[r * n + j] = fma(-f, I[p * n + j], I[r * n + j]);
      }
    }
  }
  memcpy(B, I, sizeof(double) * n * n);
  if (rc != 0) {
    dm = 0.0;
    de = 0.0;
  }
  if (mant) *mant = dm;
  if (ex) *ex = de;
  return rc;
}

int orc_sj_inverse(const orc_sj *wf, const double *cfg, int spin, double *B) {
  return orc_sj_inverse_det(wf, cfg, spin, B, NULL, NULL);
}

/* ---- Jastrow ------------------------------------------------------- */

static double pair_b(const orc_sj *wf, int i, int j) {
  return spin_of(wf, i) == spin_of(wf, j) ? wf->p[1] : wf->p[0];
}

static double jas_full(const orc_sj *wf, const double *cfg) {
  int N = wf->n_up + wf->n_dn;
  double J = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      const double *a = cfg + 3 * i, *b = cfg + 3 * j;
      double r = orc_norm3(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
      J += (pair_b(wf, i, j) * r) / (1.0 + wf->p[2] * r);
    }
  for (int i = 0; i < N; ++i)
    for (int I = 0; I < wf->n_nuc; ++I) {
      const double *a = cfg + 3 * i;
      double r = orc_norm3(a[0] - wf->nuc[3 * I], a[1] - wf->nuc[3 * I + 1], a[2] - wf->nuc[3 * I + 2]);
      double aZ = wf->p[3] * (double)wf->charge[I];
      J += -(aZ * r) / (1.0 + wf->p[4] * r);
    }
  return J;
}

/* J_i(r): sum over j != i of the pair terms with electron i at r, then the
 * e-n terms; FP64 */
static double jas_i_d(const orc_sj *wf, const double *cfg, int i, const double *r) {
  int N = wf->n_up + wf->n_dn;
  double s = 0.0;
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *x = cfg + 3 * j;
    double rij = orc_norm3(r[0] - x[0], r[1] - x[1], r[2] - x[2]);
    s = s + (pair_b(wf, i, j) * rij) / (1.0 + wf->p[2] * rij);
  }
  for (int I = 0; I < wf->n_nuc; ++I) {
    double rI = orc_norm3(r[0] - wf->nuc[3 * I], r[1] - wf->nuc[3 * I + 1], r[2] - wf->nuc[3 * I + 2]);
    double aZ = wf->p[3] * (double)wf->charge[I];
    s = s + -(aZ * rI) / (1.0 + wf->p[4] * rI);
  }
  return s;
}

/* FP32 version: differences rounded to float, float arithmetic */
static float jas_i_f(const orc_sj *wf, const double *cfg, int i, const double *r) {
  int N = wf->n_up + wf->n_dn;
  const float c = (float)wf->p[2], d = (float)wf->p[4];
  float s = 0.0f;
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *x = cfg + 3 * j;
    float d0 = (float)(r[0] - x[0]), d1 = (float)(r[1] - x[1]), d2 = (float)(r[2] - x[2]);
    float q = d0 * d0;
    q = q + d1 * d1;
    q = q + d2 * d2;
    float rij = sqrtf(q);
    float b = (float)pair_b(wf, i, j);
    s = s + (b * rij) / (1.0f + c * rij);
  }
  for (int I = 0; I < wf->n_nuc; ++I) {
    float d0 = (float)(r[0] - wf->nuc[3 * I]), d1 = (float)(r[1] - wf->nuc[3 * I + 1]),
          d2 = (float)(r[2] - wf->nuc[3 * I + 2]);
    float q = d0 * d0;
    q = q + d1 * d1;
    q = q + d2 * d2;
    float rI = sqrtf(q);
    float aZ = (float)(wf->p[3] * (double)wf->charge[I]);
    s = s + -(aZ * rI) / (1.0f + d * rI);
  }
  return s;
}

double orc_sj_value(const orc_sj *wf, const double *cfg) {
  return (lu_det(wf, cfg, 0) * lu_det(wf, cfg, 1)) * orc_exp(jas_full(wf, cfg));
}

/* ---- local energy ---------------------------------------------------- */

/* the built-in local energy; kin_out / vee_out (optional) receive
 * sum_i lap_i psi/psi and sum_{i<j} 1/r_ij (per electron, then in electron
 * order) */
static double sj_local_parts(const orc_sj *wf, const double *cfg, const double *Bup, const double *Bdn,
                             double *dlog, double *kin_out, double *vee_out) {
  int N = wf->n_up + wf->n_dn;
  const double bA = wf->p[0], bP = wf->p[1], c = wf->p[2], a = wf->p[3], d = wf->p[4];
  double kin = 0.0, V = 0.0, vee = 0.0;
  double dl[ORC_SJ_NPARAM] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < N; ++i) {
    int s = spin_of(wf, i), cc = col_of(wf, i), n = nspin(wf, s);
    const double *B = s == 0 ? Bup : Bdn;
    const double *xi = cfg + 3 * i;
    double dI[MAXNUC][3], rI[MAXNUC];
    for (int I = 0; I < wf->n_nuc; ++I) {
      for (int q = 0; q < 3; ++q) dI[I][q] = xi[q] - wf->nuc[3 * I + q];
      rI[I] = orc_norm3(dI[I][0], dI[I][1], dI[I][2]);
    }
    double chi[MAXB], gch[MAXB][3], lch[MAXB];
    for (int m = 0; m < wf->n_basis; ++m) {
      int I = wf->center[m];
      basis_vgl(kind_of(wf, m), wf->zeta[m], dI[I], rI[I], &chi[m], gch[m], &lch[m]);
    }
    double Dg[3] = {0, 0, 0}, Dl = 0.0;
    for (int k = 0; k < n; ++k) {
      const double *cf = wf->coef + (size_t)k * wf->n_basis;
      double g[3] = {0, 0, 0}, l = 0.0;
      for (int m = 0; m < wf->n_basis; ++m) {
        for (int q = 0; q < 3; ++q) g[q] = fma(cf[m], gch[m][q], g[q]);
        l = fma(cf[m], lch[m], l);
      }
      double bkc = B[k * n + cc];
      for (int q = 0; q < 3; ++q) Dg[q] = fma(g[q], bkc, Dg[q]);
      Dl = fma(l, bkc, Dl);
    }
    /* electron i's parameter derivatives are summed on their own and added
     * to the totals in electron order (the device's order) */
    double gJ[3] = {0, 0, 0}, lJ = 0.0, Vi = 0.0, Vee = 0.0, dli[ORC_SJ_NPARAM] = {0, 0, 0, 0, 0, 0};
    if (wf->opt_exp) {
      /* sum_m zeta_m d ln psi / d zeta_m over electron i's row: d chi_m /
       * d zeta_m = -r chi_m (s- and p-type STO) or -r^2 chi_m (GTO); fma
       * chains in m, then k order */
      double ez = 0.0;
      for (int k = 0; k < n; ++k) {
        const double *cf = wf->coef + (size_t)k * wf->n_basis;
        double oz = 0.0;
        for (int m = 0; m < wf->n_basis; ++m) {
          const double r = rI[wf->center[m]];
          const double rp = kind_of(wf, m) == ORC_BASIS_GTO ? r * r : r;
          oz = fma(cf[m], (-(wf->zeta[m] * rp)) * chi[m], oz);
        }
        ez = fma(oz, B[k * n + cc], ez);
      }
      dli[5] = ez;
    }
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      const double *xj = cfg + 3 * j;
      double dv[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      double r = orc_norm3(dv[0], dv[1], dv[2]);
      double b = (spin_of(wf, j) == s) ? bP : bA;
      /* u = b r / (1 + c r): u' = b iden^2, u'' = -2 b c iden^3 with
       * iden = 1 / (1 + c r); grad_i u = u' dv / r, lap_i u = u'' + 2 u' / r;
       * d u / d b = r iden, d u / d c = -b r^2 iden^2 (two divisions: 1/r
       * and iden, the device's operations) */
      double ir = 1.0 / r;
      double iden = 1.0 / (1.0 + c * r);
      double iden2 = iden * iden;
      double up = b * iden2;
      double upp = (-2.0 * b * c) * (iden2 * iden);
      double ur = up * ir;
      for (int q = 0; q < 3; ++q) gJ[q] += ur * dv[q];
      lJ += upp + 2.0 * ur;
      if (j > i) {
        Vi += ir;
        Vee += ir;
        double t = r * iden;
        if (spin_of(wf, j) == s)
          dli[1] += t;
        else
          dli[0] += t;
        dli[2] += (-b * r * r) * iden2;
      }
    }
    for (int I = 0; I < wf->n_nuc; ++I) {
      double Z = (double)wf->charge[I];
      double r = rI[I];
      /* v = -a Z r / (1 + d r), the same forms with iden = 1 / (1 + d r) */
      double ir = 1.0 / r;
      double iden = 1.0 / (1.0 + d * r);
      double iden2 = iden * iden;
      double aZ = a * Z;
      double vp = -aZ * iden2;
      double vpp = (2.0 * aZ * d) * (iden2 * iden);
      double vr = vp * ir;
      for (int q = 0; q < 3; ++q) gJ[q] += vr * dI[I][q];
      lJ += vpp + 2.0 * vr;
      Vi += -Z * ir;
      dli[3] += (-Z * r) * iden;
      dli[4] += (aZ * r * r) * iden2;
    }
    double dot = Dg[0] * gJ[0] + Dg[1] * gJ[1] + Dg[2] * gJ[2];
    double gg = gJ[0] * gJ[0] + gJ[1] * gJ[1] + gJ[2] * gJ[2];
    kin += Dl + 2.0 * dot + lJ + gg;
    V += Vi;
    vee += Vee;
    for (int q = 0; q < ORC_SJ_NPARAM; ++q) dl[q] += dli[q];
  }
  double vnn = 0.0;
  for (int I = 0; I < wf->n_nuc; ++I)
    for (int J = I + 1; J < wf->n_nuc; ++J) {
      double r = orc_norm3(wf->nuc[3 * J] - wf->nuc[3 * I], wf->nuc[3 * J + 1] - wf->nuc[3 * I + 1],
                           wf->nuc[3 * J + 2] - wf->nuc[3 * I + 2]);
      vnn += (double)(wf->charge[I] * wf->charge[J]) / r;
    }
  if (wf->hamiltonian != ORC_H_ELECTRONIC) {
    /* IonicHamiltonian: drop the e-e repulsion */
    V = 0.0;
    for (int i = 0; i < N; ++i)
      for (int I = 0; I < wf->n_nuc; ++I) {
        const double *xi = cfg + 3 * i;
        V += -(double)wf->charge[I] /
             orc_norm3(xi[0] - wf->nuc[3 * I], xi[1] - wf->nuc[3 * I + 1], xi[2] - wf->nuc[3 * I + 2]);
      }
  }
  /* d ln psi / d s = (1/s) sum_m zeta_m d ln psi / d zeta_m */
  if (wf->opt_exp) dl[5] = dl[5] / wf->p[5];
  if (dlog) memcpy(dlog, dl, sizeof dl);
  if (kin_out) *kin_out = kin;
  if (vee_out) *vee_out = vee;
  return -0.5 * kin + V + vnn;
}

double orc_sj_local_energy(const orc_sj *wf, const double *cfg, const double *Bup, const double *Bdn,
                           double *dlog) {
  return sj_local_parts(wf, cfg, Bup, Bdn, dlog, NULL, NULL);
}

static double nuclear_repulsion(const orc_sj *wf) {
  double vnn = 0.0;
  for (int I = 0; I < wf->n_nuc; ++I)
    for (int J = I + 1; J < wf->n_nuc; ++J) {
      double r = orc_norm3(wf->nuc[3 * J] - wf->nuc[3 * I], wf->nuc[3 * J + 1] - wf->nuc[3 * I + 1],
                           wf->nuc[3 * J + 2] - wf->nuc[3 * I + 2]);
      vnn += (double)(wf->charge[I] * wf->charge[J]) / r;
    }
  return vnn;
}

double orc_sj_operator(const orc_sj *wf, const double *cfg, const double *Bup, const double *Bdn,
                       const orc_operator *op) {
  const int N = wf->n_up + wf->n_dn;
  double kin, vee, ven = 0.0;
  sj_local_parts(wf, cfg, Bup, Bdn, NULL, &kin, &vee);
  for (int I = 0; I < wf->n_nuc; ++I)
    for (int i = 0; i < N; ++i) {
      const double *xi = cfg + 3 * i;
      ven -= (double)wf->charge[I] /
             orc_norm3(xi[0] - wf->nuc[3 * I], xi[1] - wf->nuc[3 * I + 1], xi[2] - wf->nuc[3 * I + 2]);
    }
  double v = 0.0;
  for (int k = 0; k < op->n_terms; ++k) {
    const orc_term *t = &op->term[k];
    double x = 0.0;
    switch (t->kind) {
      case ORC_TERM_KINETIC: x = -0.5 * kin; break;
      case ORC_TERM_IONIC_POTENTIAL: x = ven + nuclear_repulsion(wf); break;
      case ORC_TERM_ELECTRONIC_POTENTIAL: x = vee; break;
      case ORC_TERM_HARMONIC: {
        double s2 = 0.0;
        for (int i = 0; i < N; ++i)
          for (int q = 0; q < 3; ++q) {
            const double dq = cfg[3 * i + q] - t->param[1 + q];
            s2 += dq * dq;
          }
        x = 0.5 * (t->param[0] * t->param[0]) * s2;
        break;
      }
      case ORC_TERM_RADIAL_POWER: {
        const int pw = (int)t->param[0];
        for (int i = 0; i < N; ++i) {
          const double r =
              orc_norm3(cfg[3 * i] - t->param[1], cfg[3 * i + 1] - t->param[2], cfg[3 * i + 2] - t->param[3]);
          const double b = pw > 0 ? r : 1.0 / r;
          double y = b;
          for (int e = 1; e < (pw > 0 ? pw : -pw); ++e) y *= b;
          x += y;
        }
        break;
      }
      case ORC_TERM_DIPOLE:
        for (int i = 0; i < N; ++i)
          x += ((cfg[3 * i] - t->param[0]) * t->param[3] + (cfg[3 * i + 1] - t->param[1]) * t->param[4]) +
               (cfg[3 * i + 2] - t->param[2]) * t->param[5];
        break;
      default: break;
    }
    v += t->coef * x;
  }
  return v;
}

double orc_sj_local_energy_full(const orc_sj *wf, const double *cfg, double *dlog) {
  double Bu[MAXN * MAXN], Bd[MAXN * MAXN];
  orc_sj_inverse(wf, cfg, 0, Bu);
  orc_sj_inverse(wf, cfg, 1, Bd);
  return orc_sj_local_energy(wf, cfg, Bu, Bd, dlog);
}

/* ---- drift (grad_i ln psi) pieces for the diffusion proposal ---------- */

/* sum_k grad phi_k(r) Bcol[k]: basis derivatives at r, orbital gradients as
 * fma chains over the basis, then fma chains over k (FP64). */
static void det_grad(const orc_sj *wf, const double *r, int n, const double *Bcol, double g[3]) {
  double gch[MAXB][3];
  for (int m = 0; m < wf->n_basis; ++m) {
    int I = wf->center[m];
    double d[3] = {r[0] - wf->nuc[3 * I], r[1] - wf->nuc[3 * I + 1], r[2] - wf->nuc[3 * I + 2]};
    double rr = orc_norm3(d[0], d[1], d[2]);
    double v, l;
    basis_vgl(kind_of(wf, m), wf->zeta[m], d, rr, &v, gch[m], &l);
  }
  g[0] = g[1] = g[2] = 0.0;
  for (int k = 0; k < n; ++k) {
    const double *cf = wf->coef + (size_t)k * wf->n_basis;
    double gk[3] = {0, 0, 0};
    for (int m = 0; m < wf->n_basis; ++m)
      for (int q = 0; q < 3; ++q) gk[q] = fma(cf[m], gch[m][q], gk[q]);
    for (int q = 0; q < 3; ++q) g[q] = fma(gk[q], Bcol[k], g[q]);
  }
}

/* grad of J_i at electron i placed at r: pair terms j != i in order, then
 * the nuclei (FP64). */
static void jas_grad(const orc_sj *wf, const double *cfg, int i, const double *r, double g[3]) {
  int N = wf->n_up + wf->n_dn;
  const double c = wf->p[2], a = wf->p[3], d = wf->p[4];
  g[0] = g[1] = g[2] = 0.0;
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *xj = cfg + 3 * j;
    double dv[3] = {r[0] - xj[0], r[1] - xj[1], r[2] - xj[2]};
    double rr = orc_norm3(dv[0], dv[1], dv[2]);
    double b = pair_b(wf, i, j);
    double den = 1.0 + c * rr;
    double up = b / (den * den);
    for (int q = 0; q < 3; ++q) g[q] += up * dv[q] / rr;
  }
  for (int I = 0; I < wf->n_nuc; ++I) {
    double dI[3] = {r[0] - wf->nuc[3 * I], r[1] - wf->nuc[3 * I + 1], r[2] - wf->nuc[3 * I + 2]};
    double rr = orc_norm3(dI[0], dI[1], dI[2]);
    double den = 1.0 + d * rr;
    double vp = -(a * (double)wf->charge[I]) / (den * den);
    for (int q = 0; q < 3; ++q) g[q] += vp * dI[q] / rr;
  }
}

void orc_sj_drift(const orc_sj *wf, const double *cfg, int i, double *v) {
  int sp = spin_of(wf, i), c = col_of(wf, i), n = nspin(wf, sp);
  double B[MAXN * MAXN], col[MAXN], gd[3], gj[3];
  orc_sj_inverse(wf, cfg, sp, B);
  for (int k = 0; k < n; ++k) col[k] = B[k * n + c];
  det_grad(wf, cfg + 3 * i, n, col, gd);
  jas_grad(wf, cfg, i, cfg + 3 * i, gj);
  for (int q = 0; q < 3; ++q) v[q] = gd[q] + gj[q];
}

void orc_sj_move_terms(const orc_sj *wf, const double *cfg, int i, const double *xn, double *lR2,
                       double *g) {
  int sp = spin_of(wf, i), c = col_of(wf, i), n = nspin(wf, sp);
  double B[MAXN * MAXN], col[MAXN], chi[MAXB], gd[3], gj[3];
  orc_sj_inverse(wf, cfg, sp, B);
  for (int k = 0; k < n; ++k) col[k] = B[k * n + c];
  basis_at(wf, xn, chi);
  double rd = 0.0;
  for (int k = 0; k < n; ++k) rd = fma(orbital(wf, k, chi), col[k], rd);
  double dJ = jas_i_d(wf, cfg, i, xn) - jas_i_d(wf, cfg, i, cfg + 3 * i);
  *lR2 = 2.0 * log(fabs(rd)) + 2.0 * dJ;
  det_grad(wf, xn, n, col, gd);
  jas_grad(wf, cfg, i, xn, gj);
  for (int q = 0; q < 3; ++q) g[q] = gd[q] / rd + gj[q];
}

/* Accept rule of the diffusion move (metrop.rs:178-197, electron i's row):
 * Bcol[k * stride] is column c of the inverse before the move, rd the
 * determinant ratio, R2 = rd^2 exp(2 dJ), v the drift at x. */
static int diffuse_accept(const orc_sj *wf, const double *cfg, int i, const double *xn, int n,
                          const double *Bcol, int stride, double rd, double R2, const double *v,
                          double tau, double u) {
  double col[MAXN];
  for (int k = 0; k < n; ++k) col[k] = Bcol[k * stride];
  double gd[3], gj[3], vn[3];
  det_grad(wf, xn, n, col, gd);
  jas_grad(wf, cfg, i, xn, gj);
  for (int q = 0; q < 3; ++q) vn[q] = gd[q] / rd + gj[q];
  const double *x = cfg + 3 * i;
  double eh[3], el[3];
  for (int q = 0; q < 3; ++q) {
    eh[q] = (x[q] - xn[q]) - vn[q] * tau;
    el[q] = (xn[q] - x[q]) - v[q] * tau;
  }
  double nh = orc_norm3(eh[0], eh[1], eh[2]);
  double nl = orc_norm3(el[0], el[1], el[2]);
  double t_high = orc_exp(-(nh * nh) / (2.0 * tau));
  double t_low = orc_exp(-(nl * nl) / (2.0 * tau));
  if (rd < 0.0) return 0;
  return rust_min1((t_high * R2) / t_low) > u;
}

/* ---- runners ---------------------------------------------------------- */

typedef struct {
  const orc_sj *wf;
  const orc_sj_run_params *rp;
  orc_sj_run_out *out;
  uint32_t w0, w1;
  int faithful;
  double sum_e, sum_e2, sum_bm, sum_bm2;
  uint64_t n_samples, n_blocks, n_accepted, n_moves;
  double obs_sum[8], obs_sum2[8];
  double max_ld_drift;
  uint64_t n_singular;
} sj_task;

/* Re-inverts both spin blocks (the device's refresh).  With ld (the
 * walker's carried determinants, mant/ex per spin): ld_mode 1 replaces them
 * by the re-inversion's; 2 first records the drift of the carried log|det|
 * in *drift (non-singular blocks whose carried mantissa is not 0).  Singular
 * blocks are counted in *n_singular (may be NULL). */
static void refresh_ld(const orc_sj *wf, const double *cfg, int prec, double *Bu, double *Bd, float *Buf, float *Bdf,
                       double *ld, int ld_mode, double *drift, uint64_t *n_singular) {
  double m[2], e[2];
  const int r0 = orc_sj_inverse_det(wf, cfg, 0, Bu, &m[0], &e[0]);
  const int r1 = orc_sj_inverse_det(wf, cfg, 1, Bd, &m[1], &e[1]);
  if (n_singular) *n_singular += (uint64_t)(r0 != 0) + (uint64_t)(r1 != 0);
  const int bad[2] = {r0 != 0, r1 != 0};
  for (int q = 0; q < 2 && ld; ++q) {
    if (ld_mode == 2 && !bad[q] && ld[2 * q] != 0.0) {
      const double carried = orc_sj_logdet_value(ld[2 * q], ld[2 * q + 1]);
      const double dv = fabs(carried - orc_sj_logdet_value(m[q], e[q]));
      *drift = fmax(*drift, dv);
    }
    if (ld_mode) {
      ld[2 * q] = m[q];
      ld[2 * q + 1] = e[q];
    }
  }
  if (prec == ORC_PREC_MIXED) {
    for (int k = 0; k < wf->n_up * wf->n_up; ++k) Buf[k] = (float)Bu[k];
    for (int k = 0; k < wf->n_dn * wf->n_dn; ++k) Bdf[k] = (float)Bd[k];
  }
}

static void refresh(const orc_sj *wf, const double *cfg, int prec, double *Bu, double *Bd, float *Buf, float *Bdf) {
  refresh_ld(wf, cfg, prec, Bu, Bd, Buf, Bdf, NULL, 0, NULL, NULL);
}

/* One sweep of walker w over all electrons (the body of Runner::run's
 * move_state loop, samplers.rs:106-117): draws from *pos on, global move
 * counter *m; returns the accept mask. */
static uint64_t sj_sweep_once(sj_task *t, uint32_t w, double *cfg, double *Bu, double *Bd, float *Buf, float *Bdf,
                              double *ld, uint64_t *pos_io, uint64_t *m_io) {
  const orc_sj *wf = t->wf;
  const orc_sj_run_params *rp = t->rp;
  const int N = wf->n_up + wf->n_dn;
  const int prec = rp->precision;
  const double half = 0.5 * rp->box_side;
  double chi[MAXB];
  uint64_t pos = *pos_io, m = *m_io;
  uint64_t mask = 0;
  for (int i = 0; i < N; ++i) {
    double xn[3], u;
    double vold[3] = {0, 0, 0}; /* drift at x (diffusion) */
    const int diffuse = !t->faithful && rp->proposal == ORC_SJ_DIFFUSE;
    const int learned = !t->faithful && rp->proposal == ORC_SJ_LEARNED;
    double lq_ratio = 0.0;
    if (learned) {
      uint64_t dr[6];
      for (int k = 0; k < 6; ++k) dr[k] = orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++);
      double z[4];
      orc_box_muller(dr[0], dr[1], &z[0], &z[1]);
      orc_box_muller(dr[2], dr[3], &z[2], &z[3]);
      u = orc_u64_standard(dr[4]);
      double f1[ORC_MLP_NIN], f2[ORC_MLP_NIN], o1[ORC_MLP_NOUT], o2[ORC_MLP_NOUT];
      orc_mlp_features(wf, cfg, i, cfg + 3 * i, f1);
      orc_mlp_forward(rp->mlp_w, f1, rp->mlp_emulate_bf16, o1, NULL);
      double ls = orc_mlp_log_sigma(o1[3]), sg = exp(ls), z2 = 0.0;
      for (int q = 0; q < 3; ++q) {
        xn[q] = (cfg[3 * i + q] + o1[q]) + sg * z[q];
        z2 += z[q] * z[q];
      }
      double cfg2[3 * MAXE];
      memcpy(cfg2, cfg, sizeof(double) * 3 * N);
      memcpy(cfg2 + 3 * i, xn, sizeof xn);
      orc_mlp_features(wf, cfg2, i, xn, f2);
      orc_mlp_forward(rp->mlp_w, f2, rp->mlp_emulate_bf16, o2, NULL);
      double ls2 = orc_mlp_log_sigma(o2[3]), sg2 = exp(ls2), e2 = 0.0;
      for (int q = 0; q < 3; ++q) {
        double e = (cfg[3 * i + q] - xn[q]) - o2[q];
        e2 += e * e;
      }
      lq_ratio = (-3.0 * ls2 - e2 / (2.0 * sg2 * sg2)) - (-3.0 * ls - 0.5 * z2);
      if (rp->mlp_drop_q_ratio) lq_ratio = 0.0;
    } else if (!diffuse) {
      double dd[3];
      for (int k = 0; k < 3; ++k)
        dd[k] = orc_u64_range(orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++), -half, half);
      u = orc_u64_standard(orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++));
      for (int k = 0; k < 3; ++k) xn[k] = cfg[3 * i + k] + dd[k];
    } else {
      uint64_t dr[6];
      for (int k = 0; k < 6; ++k) dr[k] = orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, pos++);
      double z[4];
      orc_box_muller(dr[0], dr[1], &z[0], &z[1]);
      orc_box_muller(dr[2], dr[3], &z[2], &z[3]);
      u = orc_u64_standard(dr[4]);
      int sp = spin_of(wf, i), c = col_of(wf, i), n = nspin(wf, sp);
      double Bcol[MAXN];
      for (int k = 0; k < n; ++k)
        Bcol[k] = prec == ORC_PREC_FP64 ? (sp == 0 ? Bu : Bd)[k * n + c] : (double)(sp == 0 ? Buf : Bdf)[k * n + c];
      double gd[3], gj[3];
      det_grad(wf, cfg + 3 * i, n, Bcol, gd);
      jas_grad(wf, cfg, i, cfg + 3 * i, gj);
      const double st = sqrt(rp->tau);
      for (int q = 0; q < 3; ++q) {
        vold[q] = gd[q] + gj[q];
        xn[q] = (cfg[3 * i + q] + vold[q] * rp->tau) + st * z[q];
      }
    }
    int acc;
    t->n_moves++;
    if (t->faithful) {
      double prop[3 * MAXE];
      memcpy(prop, cfg, sizeof(double) * 3 * N);
      memcpy(prop + 3 * i, xn, sizeof xn);
      double pn = orc_sj_value(wf, prop);
      double po = orc_sj_value(wf, cfg);
      acc = rust_min1((pn * pn) / (po * po)) > u;
    } else {
      int sp = spin_of(wf, i), c = col_of(wf, i), n = nspin(wf, sp);
      basis_at(wf, xn, chi);
      if (prec == ORC_PREC_FP64) {
        double *B = sp == 0 ? Bu : Bd;
        double phin[MAXN];
        for (int k = 0; k < n; ++k) phin[k] = orbital(wf, k, chi);
        double rd = 0.0;
        for (int k = 0; k < n; ++k) rd = fma(phin[k], B[k * n + c], rd);
        double Jn = jas_i_d(wf, cfg, i, xn);
        double Jo = jas_i_d(wf, cfg, i, cfg + 3 * i);
        double dJ = Jn - Jo;
        double R2 = (rd * rd) * orc_exp(2.0 * dJ);
        acc = diffuse   ? diffuse_accept(wf, cfg, i, xn, n, B + c, n, rd, R2, vold, rp->tau, u)
              : learned ? rust_min1(R2 * exp(lq_ratio)) > u
                        : rust_min1(R2) > u;
        if (acc) orc_sj_logdet_mul(&ld[2 * sp], &ld[2 * sp + 1], rd);
        if (acc) {
          double inv = 1.0 / rd;
          double f[MAXN];
          for (int l = 0; l < n; ++l) {
            if (l == c) continue;
            double v = 0.0;
            for (int k = 0; k < n; ++k) v = fma(phin[k], B[k * n + l], v);
            f[l] = v * inv;
          }
          for (int k = 0; k < n; ++k)
            for (int l = 0; l < n; ++l)
              if (l != c) B[k * n + l] = fma(-B[k * n + c], f[l], B[k * n + l]);
          for (int k = 0; k < n; ++k) B[k * n + c] = B[k * n + c] * inv;
        }
      } else {
        float *B = sp == 0 ? Buf : Bdf;
        float phin[MAXN];
        for (int k = 0; k < n; ++k) phin[k] = (float)orbital(wf, k, chi);
        float rd = 0.0f;
        for (int k = 0; k < n; ++k) rd = fmaf(phin[k], B[k * n + c], rd);
        float Jn = jas_i_f(wf, cfg, i, xn);
        float Jo = jas_i_f(wf, cfg, i, cfg + 3 * i);
        float dJ = Jn - Jo;
        double rdd = (double)rd;
        double R2 = (rdd * rdd) * orc_exp(2.0 * (double)dJ);
        if (diffuse) {
          double Bcol[MAXN];
          for (int k = 0; k < n; ++k) Bcol[k] = (double)B[k * n + c];
          acc = diffuse_accept(wf, cfg, i, xn, n, Bcol, 1, rdd, R2, vold, rp->tau, u);
        } else if (learned) {
          acc = rust_min1(R2 * exp(lq_ratio)) > u;
        } else {
          acc = rust_min1(R2) > u;
        }
        if (acc) orc_sj_logdet_mul(&ld[2 * sp], &ld[2 * sp + 1], rdd);
        if (acc) {
          float inv = 1.0f / rd;
          float f[MAXN];
          for (int l = 0; l < n; ++l) {
            if (l == c) continue;
            float v = 0.0f;
            for (int k = 0; k < n; ++k) v = fmaf(phin[k], B[k * n + l], v);
            f[l] = v * inv;
          }
          for (int k = 0; k < n; ++k)
            for (int l = 0; l < n; ++l)
              if (l != c) B[k * n + l] = fmaf(-B[k * n + c], f[l], B[k * n + l]);
          for (int k = 0; k < n; ++k) B[k * n + c] = B[k * n + c] * inv;
        }
      }
    }
    if (acc) {
      memcpy(cfg + 3 * i, xn, sizeof xn);
      mask |= (uint64_t)1 << i;
      t->n_accepted++;
    }
    ++m;
    if (!t->faithful && rp->lu_period && m % rp->lu_period == 0)
      refresh_ld(wf, cfg, prec, Bu, Bd, Buf, Bdf, ld, 2, &t->max_ld_drift, &t->n_singular);
  }
  *pos_io = pos;
  *m_io = m;
  return mask;
}

static void sj_walker(sj_task *t, uint32_t w) {
  const orc_sj *wf = t->wf;
  const orc_sj_run_params *rp = t->rp;
  orc_sj_run_out *out = t->out;
  const int N = wf->n_up + wf->n_dn;
  const int prec = rp->precision;
  uint32_t blocks = rp->steps / rp->block_size;
  uint64_t nsw = (uint64_t)(blocks - 1) * rp->block_size;
  double cfg[3 * MAXE];
  uint64_t pos;
  if (rp->init_from_stream) {
    for (int k = 0; k < 3 * N; ++k)
      cfg[k] = orc_u64_range(orc_stream_u64(rp->seed, rp->sys, w, rp->epoch, (uint64_t)k), -1.0, 1.0);
    pos = 3 * (uint64_t)N;
  } else {
    memcpy(cfg, out->cfg + (size_t)w * 3 * N, sizeof(double) * 3 * N);
    pos = rp->pos0;
  }
  double Bu[MAXN * MAXN], Bd[MAXN * MAXN];
  float Buf[MAXN * MAXN], Bdf[MAXN * MAXN];
  double ld[4] = {1.0, 0.0, 1.0, 0.0};
  if (!t->faithful) refresh_ld(wf, cfg, prec, Bu, Bd, Buf, Bdf, ld, 1, NULL, &t->n_singular);
  uint64_t m = rp->move0;
  uint64_t sidx = 0;
  for (uint32_t block = 0; block < blocks; ++block) {
    double bsum = 0.0;
    for (uint32_t s = 0; s < rp->block_size; ++s) {
      const uint64_t mask = sj_sweep_once(t, w, cfg, Bu, Bd, Buf, Bdf, ld, &pos, &m);
      if (out->accept_mask)
        out->accept_mask[(size_t)w * blocks * rp->block_size + (size_t)block * rp->block_size + s] = mask;
      if (block > 0) {
        double dl[ORC_SJ_NPARAM];
        double el;
        if (t->faithful) {
          el = orc_sj_local_energy_full(wf, cfg, dl);
        } else if (prec == ORC_PREC_FP64) {
          el = orc_sj_local_energy(wf, cfg, Bu, Bd, dl);
        } else {
          double bu[MAXN * MAXN], bd[MAXN * MAXN];
          for (int k = 0; k < wf->n_up * wf->n_up; ++k) bu[k] = (double)Buf[k];
          for (int k = 0; k < wf->n_dn * wf->n_dn; ++k) bd[k] = (double)Bdf[k];
          el = orc_sj_local_energy(wf, cfg, bu, bd, dl);
        }
        if (!t->faithful && (rp->hamiltonian || rp->n_obs)) {
          double bu[MAXN * MAXN], bd[MAXN * MAXN];
          for (int k = 0; k < wf->n_up * wf->n_up; ++k) bu[k] = prec == ORC_PREC_FP64 ? Bu[k] : (double)Buf[k];
          for (int k = 0; k < wf->n_dn * wf->n_dn; ++k) bd[k] = prec == ORC_PREC_FP64 ? Bd[k] : (double)Bdf[k];
          if (rp->hamiltonian) el = orc_sj_operator(wf, cfg, bu, bd, rp->hamiltonian);
          for (int k = 0; k < rp->n_obs; ++k) {
            const double v = orc_sj_operator(wf, cfg, bu, bd, &rp->obs[k]);
            t->obs_sum[k] += v;
            t->obs_sum2[k] += v * v;
          }
        }
        size_t si = (size_t)w * nsw + sidx;
        if (out->e_local) out->e_local[si] = el;
        if (out->dlog) memcpy(out->dlog + si * ORC_SJ_NPARAM, dl, sizeof dl);
        if (out->psi && !t->faithful) out->psi[si] = orc_sj_compat_psi(wf, ld, dl);
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
  if (out->cfg) memcpy(out->cfg + (size_t)w * 3 * N, cfg, sizeof(double) * 3 * N);
  if (out->ld && !t->faithful) memcpy(out->ld + (size_t)w * 4, ld, sizeof ld);
}

static void *sj_task_main(void *arg) {
  sj_task *t = (sj_task *)arg;
  for (uint32_t w = t->w0; w < t->w1; ++w) sj_walker(t, w);
  return NULL;
}

static int sj_run(const orc_sj *wf, const orc_sj_run_params *rp, uint32_t w0, uint32_t w1,
                  int n_threads, orc_sj_run_out *out, int faithful) {
  int N = wf->n_up + wf->n_dn;
  if (rp->block_size == 0 || rp->steps < 2 * rp->block_size) return -1;
  if (N > MAXE || wf->n_up > MAXN || wf->n_dn > MAXN || wf->n_nuc > MAXNUC || wf->n_basis > MAXB)
    return -1;
  if (out->accept_mask && N > 64) return -1;
  if (!rp->init_from_stream && !out->cfg) return -1;
  if (rp->n_obs < 0 || rp->n_obs > 8) return -1;
  if (n_threads < 1) n_threads = 1;
  uint32_t nw = w1 - w0;
  if ((uint32_t)n_threads > nw) n_threads = nw ? (int)nw : 1;
  sj_task *tasks = (sj_task *)calloc((size_t)n_threads, sizeof(sj_task));
  pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
  for (int i = 0; i < n_threads; ++i) {
    tasks[i].wf = wf;
    tasks[i].rp = rp;
    tasks[i].out = out;
    tasks[i].faithful = faithful;
    tasks[i].w0 = w0 + (uint32_t)((uint64_t)nw * i / n_threads);
    tasks[i].w1 = w0 + (uint32_t)((uint64_t)nw * (i + 1) / n_threads);
  }
  if (n_threads == 1)
    sj_task_main(&tasks[0]);
  else {
    for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, sj_task_main, &tasks[i]);
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
    out->max_ld_drift = fmax(out->max_ld_drift, tasks[i].max_ld_drift);
    out->n_singular += tasks[i].n_singular;
    for (int k = 0; k < rp->n_obs; ++k) {
      if (out->obs_sum) out->obs_sum[k] += tasks[i].obs_sum[k];
      if (out->obs_sum2) out->obs_sum2[k] += tasks[i].obs_sum2[k];
    }
  }
  free(tasks);
  free(th);
  return 0;
}

int orc_sj_run_parity(const orc_sj *wf, const orc_sj_run_params *rp, uint32_t w0, uint32_t w1,
                      int n_threads, orc_sj_run_out *out) {
  return sj_run(wf, rp, w0, w1, n_threads, out, 0);
}

int orc_sj_run_faithful(const orc_sj *wf, const orc_sj_run_params *rp, uint32_t w0, uint32_t w1,
                        int n_threads, orc_sj_run_out *out) {
  return sj_run(wf, rp, w0, w1, n_threads, out, 1);
}

/* ------------------------------------------------------------------ */
/* Diffusion Monte Carlo                                               */
/* ------------------------------------------------------------------ */

static uint64_t uniform_below(uint64_t d, uint64_t n) {
  return (uint64_t)(((unsigned __int128)d * n) >> 64);
}

/* Rust's `f64 as u32`: truncation, saturating, NaN -> 0 */
static uint32_t rust_as_u32(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

/* Rust's `f64 as usize` (64-bit) */
static uint64_t rust_as_usize(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)x;
}

/* Sums over walkers in the device's fixed order (the reference's
 * sequential fold, `walkers.iter().fold`, differs from it in the last bits;
 * its order is ndarray's/std's and not pinned): blocks of 4096 walkers, in
 * each block thread t of 256 adds elements t, t + 256, ... in order, then a
 * pairwise tree over the threads (offsets 128, 64, ..., 1); the block sums
 * are reduced the same way (thread t adds blocks t, t + 256, ...; tree). */
static double tree256(double *sh) {
  for (int off = 128; off > 0; off >>= 1)
    for (int t = 0; t < off; ++t) sh[t] += sh[t + off];
  return sh[0];
}

double orc_device_sum(const double *x, size_t n) {
  const size_t chunk = 4096, nb = (n + chunk - 1) / chunk;
  double *part = (double *)malloc(sizeof(double) * (nb ? nb : 1));
  double sh[256];
  for (size_t b = 0; b < nb; ++b) {
    for (int t = 0; t < 256; ++t) {
      double s = 0.0;
      for (int k = 0; k < 16; ++k) {
        const size_t i = b * chunk + (size_t)k * 256 + (size_t)t;
        if (i < n) s += x[i];
      }
      sh[t] = s;
    }
    part[b] = tree256(sh);
  }
  for (int t = 0; t < 256; ++t) {
    double s = 0.0;
    for (size_t i = (size_t)t; i < nb; i += 256) s += part[i];
    sh[t] = s;
  }
  const double r = tree256(sh);
  free(part);
  return r;
}

int orc_branch_sr(const double *w, uint32_t W, uint64_t seed, uint32_t sys, uint64_t step, uint32_t *src,
                  double *w_out) {
  double mx = 0.0;
  for (uint32_t i = 0; i < W; ++i) mx = fmax(mx, w[i]); /* f64::max ignores NaN; max is order-free */
  const double tot = orc_device_sum(w, W);
  const double global_weight = tot / (double)W;
  const double norm = (double)W / mx;
  uint64_t *cum = (uint64_t *)malloc(sizeof(uint64_t) * (W ? W : 1));
  uint64_t run = 0;
  for (uint32_t i = 0; i < W; ++i) {
    run += rust_as_u32(w[i] * norm);
    cum[i] = run;
  }
  if (run == 0) {
    free(cum);
    return -1;
  }
  for (uint32_t j = 0; j < W; ++j) {
    const uint64_t u = uniform_below(orc_stream_u64(seed, sys, (uint32_t)step, ORC_DMC_BRANCH_EPOCH, j), run);
    /* first k with cum[k] > u */
    uint32_t lo = 0, hi = W - 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (cum[mid] > u) hi = mid;
      else lo = mid + 1;
    }
    src[j] = lo;
    w_out[j] = global_weight;
  }
  free(cum);
  return 0;
}

void orc_branch_simple(const double *w, uint32_t W, uint64_t seed, uint32_t sys, uint64_t step, uint32_t *src,
                       double *w_out) {
  /* new_walkers = walkers minus to_kill (in order) ++ to_birth */
  uint32_t *list = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (W ? W : 1));
  uint32_t *births = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (W ? W : 1));
  uint32_t M = 0, nb = 0;
  for (uint32_t i = 0; i < W; ++i) {
    const double u = orc_u64_standard(orc_stream_u64(seed, sys, (uint32_t)step, ORC_DMC_BRANCH_EPOCH, i));
    uint64_t n = rust_as_usize(w[i] + u);
    if (n > 3) n = 3;
    if (n == 0) continue;
    list[M++] = i;
    for (uint64_t c = 0; c + 1 < n; ++c) births[nb++] = i;
  }
  for (uint32_t b = 0; b < nb; ++b) list[M++] = births[b];
  if (W > M) {
    const uint32_t excess = W - M;
    for (uint32_t k = 0; k < excess; ++k)
      list[M + k] = (uint32_t)uniform_below(
          orc_stream_u64(seed, sys, (uint32_t)step, ORC_DMC_BRANCH_EPOCH, (uint64_t)W + k), W);
    M = W;
  } else {
    const uint32_t excess = M - W;
    for (uint32_t k = 0; k < excess; ++k) {
      const uint32_t r = (uint32_t)uniform_below(
          orc_stream_u64(seed, sys, (uint32_t)step, ORC_DMC_BRANCH_EPOCH, (uint64_t)W + k), M);
      memmove(list + r, list + r + 1, sizeof(uint32_t) * (M - r - 1)); /* Vec::remove */
      --M;
    }
  }
  for (uint32_t j = 0; j < W; ++j) {
    src[j] = list[j];
    w_out[j] = w[list[j]];
  }
  free(list);
  free(births);
}

static int dmc_nb(const orc_sj *wf) { return wf->n_up * wf->n_up + wf->n_dn * wf->n_dn; }

static void dmc_task(sj_task *t, const orc_sj *wf, const orc_dmc_params *p, orc_sj_run_params *rp) {
  memset(t, 0, sizeof(*t));
  memset(rp, 0, sizeof(*rp));
  rp->seed = p->seed;
  rp->sys = p->sys;
  rp->epoch = p->epoch;
  rp->precision = p->precision;
  rp->lu_period = p->lu_period;
  rp->proposal = ORC_SJ_DIFFUSE;
  rp->tau = p->tau;
  t->wf = wf;
  t->rp = rp;
}

/* the walker's inverses in the working precision (as the device stores
 * them) and its local energy from them */
static double dmc_energy(const orc_sj *wf, const orc_dmc_params *p, const double *cfg, const double *B) {
  double dl[ORC_SJ_NPARAM];
  const int nu = wf->n_up;
  if (p->hamiltonian) return orc_sj_operator(wf, cfg, B, B + nu * nu, p->hamiltonian);
  return orc_sj_local_energy(wf, cfg, B, B + nu * nu, dl);
}

int orc_dmc_init(const orc_sj *wf, const orc_dmc_params *p, double reference_energy, orc_dmc_state *st) {
  const int N = wf->n_up + wf->n_dn, nb = dmc_nb(wf);
  if (N > MAXE || wf->n_up > MAXN || wf->n_dn > MAXN || p->n_walkers == 0) return -1;
  double cfg[3 * MAXE];
  for (int k = 0; 2 * k < 3 * N; ++k) {
    double z0, z1;
    orc_box_muller(orc_stream_u64(p->seed, p->sys, 0xFFFFFFFFu, ORC_DMC_INIT_EPOCH, 2 * (uint64_t)k),
                   orc_stream_u64(p->seed, p->sys, 0xFFFFFFFFu, ORC_DMC_INIT_EPOCH, 2 * (uint64_t)k + 1), &z0, &z1);
    cfg[2 * k] = z0;
    if (2 * k + 1 < 3 * N) cfg[2 * k + 1] = z1;
  }
  double Bu[MAXN * MAXN], Bd[MAXN * MAXN];
  float Buf[MAXN * MAXN], Bdf[MAXN * MAXN];
  refresh(wf, cfg, p->precision, Bu, Bd, Buf, Bdf);
  double B[2 * MAXN * MAXN];
  for (int k = 0; k < wf->n_up * wf->n_up; ++k)
    B[k] = p->precision == ORC_PREC_FP64 ? Bu[k] : (double)Buf[k];
  for (int k = 0; k < wf->n_dn * wf->n_dn; ++k)
    B[wf->n_up * wf->n_up + k] = p->precision == ORC_PREC_FP64 ? Bd[k] : (double)Bdf[k];
  const double e0 = dmc_energy(wf, p, cfg, B);
  for (uint32_t w = 0; w < p->n_walkers; ++w) {
    memcpy(st->cfg + (size_t)w * 3 * N, cfg, sizeof(double) * 3 * N);
    memcpy(st->B + (size_t)w * nb, B, sizeof(double) * nb);
    st->weight[w] = 1.0;
    st->e_local[w] = e0;
  }
  st->reference_energy = reference_energy;
  st->steps = 0;
  st->draw = 0;
  st->moves = 0;
  return 0;
}

int orc_dmc_diffuse(const orc_sj *wf, const orc_dmc_params *p, orc_dmc_state *st, double time_step,
                    uint32_t num_iterations, uint32_t block_size, uint32_t num_eq_blocks, double *energies,
                    double *errors, double *step_energy) {
  const int N = wf->n_up + wf->n_dn, nb = dmc_nb(wf), nu = wf->n_up, nd = wf->n_dn;
  const uint32_t W = p->n_walkers;
  if (block_size == 0) return -1;
  sj_task t;
  orc_sj_run_params rp;
  dmc_task(&t, wf, p, &rp);
  double *cfg2 = (double *)malloc(sizeof(double) * W * 3 * N);
  double *B2 = (double *)malloc(sizeof(double) * W * (nb ? nb : 1));
  double *e2 = (double *)malloc(sizeof(double) * W);
  double *w2 = (double *)malloc(sizeof(double) * W);
  uint32_t *src = (uint32_t *)malloc(sizeof(uint32_t) * W);
  double *eblock = (double *)malloc(sizeof(double) * block_size);
  const uint32_t blocks = num_iterations / block_size;
  int n_out = 0;
  double *vars = (double *)malloc(sizeof(double) * (blocks ? blocks : 1));
  double *we = (double *)malloc(sizeof(double) * W), *wo = (double *)malloc(sizeof(double) * W);
  uint64_t it = 0;
  int rc = 0;
  for (uint32_t block_nr = 0; block_nr < blocks && rc == 0; ++block_nr) {
    for (uint32_t j = 0; j < block_size; ++j, ++it) {
      for (uint32_t w = 0; w < W; ++w) {
        double *c = st->cfg + (size_t)w * 3 * N;
        double *B = st->B + (size_t)w * nb;
        double Bu[MAXN * MAXN], Bd[MAXN * MAXN];
        float Buf[MAXN * MAXN], Bdf[MAXN * MAXN];
        for (int k = 0; k < nu * nu; ++k) {
          Bu[k] = B[k];
          Buf[k] = (float)B[k];
        }
        for (int k = 0; k < nd * nd; ++k) {
          Bd[k] = B[nu * nu + k];
          Bdf[k] = (float)B[nu * nu + k];
        }
        const double local_e = st->e_local[w];
        uint64_t pos = st->draw, m = st->moves;
        double ld_unused[4] = {1.0, 0.0, 1.0, 0.0}; /* DMC steps carry no determinants */
        sj_sweep_once(&t, w, c, Bu, Bd, Buf, Bdf, ld_unused, &pos, &m);
        for (int k = 0; k < nu * nu; ++k) B[k] = p->precision == ORC_PREC_FP64 ? Bu[k] : (double)Buf[k];
        for (int k = 0; k < nd * nd; ++k) B[nu * nu + k] = p->precision == ORC_PREC_FP64 ? Bd[k] : (double)Bdf[k];
        we[w] = st->weight[w] * local_e;
        wo[w] = st->weight[w];
        const double local_e_new = dmc_energy(wf, p, c, B);
        st->weight[w] *= orc_exp(-time_step * ((local_e + local_e_new) / 2.0 - st->reference_energy));
        st->e_local[w] = local_e_new;
      }
      st->draw += 6 * (uint64_t)N;
      st->moves += (uint64_t)N;
      const double ensemble = orc_device_sum(we, W) / orc_device_sum(wo, W);
      eblock[j] = ensemble;
      if (step_energy) step_energy[it] = ensemble;
      if (p->brancher == ORC_BRANCH_SR) {
        if (orc_branch_sr(st->weight, W, p->seed, p->sys, st->steps, src, w2) != 0) {
          rc = -1;
          break;
        }
      } else {
        orc_branch_simple(st->weight, W, p->seed, p->sys, st->steps, src, w2);
      }
      for (uint32_t k = 0; k < W; ++k) {
        memcpy(cfg2 + (size_t)k * 3 * N, st->cfg + (size_t)src[k] * 3 * N, sizeof(double) * 3 * N);
        memcpy(B2 + (size_t)k * nb, st->B + (size_t)src[k] * nb, sizeof(double) * nb);
        e2[k] = st->e_local[src[k]];
      }
      memcpy(st->cfg, cfg2, sizeof(double) * W * 3 * N);
      memcpy(st->B, B2, sizeof(double) * W * nb);
      memcpy(st->e_local, e2, sizeof(double) * W);
      memcpy(st->weight, w2, sizeof(double) * W);
      st->steps += 1;
    }
    if (rc) break;
    /* update_energies (dmc.rs:155-202) */
    if (block_nr == num_eq_blocks) {
      double energy = 0.0;
      for (uint32_t j = 0; j < block_size; ++j) energy += eblock[j];
      energy /= (double)block_size;
      st->reference_energy = (st->reference_energy + energy) / 2.0;
      energies[n_out] = energy;
      vars[n_out] = 0.0;
      ++n_out;
    }
    if (block_nr > num_eq_blocks) {
      double energy = 0.0;
      for (uint32_t j = 0; j < block_size; ++j) energy += eblock[j];
      energy /= (double)block_size;
      const double prev = energies[n_out - 1];
      const double k = (double)(block_nr - num_eq_blocks);
      energies[n_out] = prev + (energy - prev) / k;
      st->reference_energy = (st->reference_energy + energies[n_out]) / 2.0;
      vars[n_out] = vars[n_out - 1] + ((energy - prev) * (energy - energies[n_out]) - vars[n_out - 1]) / k;
      ++n_out;
    }
  }
  for (int i = 0; i < n_out; ++i) errors[i] = sqrt(vars[i] / (double)(i + 1));
  free(cfg2);
  free(B2);
  free(e2);
  free(w2);
  free(src);
  free(eblock);
  free(vars);
  free(we);
  free(wo);
  return rc ? -1 : n_out;
}
