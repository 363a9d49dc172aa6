/*
 * CPU oracle (test infrastructure only): the two-electron wavefunctions of
 * the He-singlet and H2-LCAO workloads, the reference's potentials and the
 * generic local-energy composition.
 *
 * The reference ships no wavefunction for these systems (README.md:31-33);
 * the user implementations these follow are
 * examples/helium_atom_singlet.rs:60-118 (product STO, its gradient,
 * laplacian and parameter gradient) and examples/hydrogen_molecule.rs:33-168
 * (LCAO STO sums), with the e-e Pade factor of theory/jastrow.tex:17-30 at
 * kappa -> 0 and no polynomial terms.
 *
 * value() is spelled once here and once on the device with the same
 * operation order, so both round identically; the derivatives below use
 * the product rule on psi (the reference's user-code style) and are
 * checked against finite differences and closed forms in tests/.
 */
#include <math.h>

#include "oracle.h"

/* Pade e-e term u(r) = b r/(1 + c r) and its r-derivatives. */
static double pade_u(double b, double c, double r) {
  return (b * r) / (1.0 + c * r);
}

/* phi(r) = sum_I exp(-zeta |r - R_I|) for the LCAO orbital; fills the
 * per-nucleus distances and exponentials. */
static double lcao_phi(const orc_wf2 *wf, const double *x, double *rI,
                       double *eI) {
  double zeta = wf->p[0];
  double phi = 0.0;
  for (int I = 0; I < wf->n_nuc; ++I) {
    rI[I] = orc_norm3(x[0] - wf->nuc[I][0], x[1] - wf->nuc[I][1],
                      x[2] - wf->nuc[I][2]);
    eI[I] = orc_exp(-zeta * rI[I]);
    phi = (I == 0) ? eI[I] : phi + eI[I];
  }
  return phi;
}

double orc_wf2_value(const orc_wf2 *wf, const double cfg[6]) {
  double zeta = wf->p[0], b = wf->p[1], c = wf->p[2];
  double r12 = orc_norm3(cfg[0] - cfg[3], cfg[1] - cfg[4], cfg[2] - cfg[5]);
  if (wf->kind == ORC_WF_HELIUM_PADE) {
    const double *R = wf->nuc[0];
    double r1 = orc_norm3(cfg[0] - R[0], cfg[1] - R[1], cfg[2] - R[2]);
    double r2 = orc_norm3(cfg[3] - R[0], cfg[4] - R[1], cfg[5] - R[2]);
    double arg = -zeta * (r1 + r2) + pade_u(b, c, r12);
    return orc_exp(arg);
  }
  double rI[ORC_MAX_NUC], eI[ORC_MAX_NUC];
  double phi1 = lcao_phi(wf, cfg, rI, eI);
  double phi2 = lcao_phi(wf, cfg + 3, rI, eI);
  double jas = orc_exp(pade_u(b, c, r12));
  return (phi1 * phi2) * jas;
}

/* Orbital factor A = phi(r1) phi(r2) pieces for one electron: value,
 * gradient and laplacian of phi at x. For the helium kind phi = e^{-zeta r}
 * about nucleus 0. */
static void orbital_vgl(const orc_wf2 *wf, const double *x, double *phi,
                        double g[3], double *lap) {
  double zeta = wf->p[0];
  int nn = (wf->kind == ORC_WF_HELIUM_PADE) ? 1 : wf->n_nuc;
  *phi = 0.0;
  g[0] = g[1] = g[2] = 0.0;
  *lap = 0.0;
  for (int I = 0; I < nn; ++I) {
    double d[3] = {x[0] - wf->nuc[I][0], x[1] - wf->nuc[I][1],
                   x[2] - wf->nuc[I][2]};
    double r = orc_norm3(d[0], d[1], d[2]);
    double e = orc_exp(-zeta * r);
    *phi += e;
    /* grad e^{-zeta r} = -zeta e r_hat; lap = e (zeta^2 - 2 zeta / r)
     * (examples/helium_atom_singlet.rs:80,96) */
    for (int k = 0; k < 3; ++k) g[k] += -zeta / r * e * d[k];
    *lap += zeta / r * e * (zeta * r - 2.0);
  }
}

/* Jastrow factor B = exp(u(r12)) with grad_1 B = u' r12_hat B,
 * grad_2 B = -grad_1 B, lap_i B = B (u'' + 2u'/r12 + u'^2). */
static void jastrow_vgl(const orc_wf2 *wf, const double cfg[6], double *B,
                        double g1[3], double *lapi) {
  double b = wf->p[1], c = wf->p[2];
  double d[3] = {cfg[0] - cfg[3], cfg[1] - cfg[4], cfg[2] - cfg[5]};
  double r = orc_norm3(d[0], d[1], d[2]);
  double den = 1.0 + c * r;
  double up = b / (den * den);
  double upp = -2.0 * b * c / (den * den * den);
  *B = orc_exp(pade_u(b, c, r));
  for (int k = 0; k < 3; ++k) g1[k] = up * d[k] / r * (*B);
  *lapi = (*B) * (upp + 2.0 * up / r + up * up);
}

void orc_wf2_gradient(const orc_wf2 *wf, const double cfg[6], double g[6]) {
  double phi1, phi2, l1, l2, ga1[3], ga2[3], B, gb1[3], lb;
  orbital_vgl(wf, cfg, &phi1, ga1, &l1);
  orbital_vgl(wf, cfg + 3, &phi2, ga2, &l2);
  jastrow_vgl(wf, cfg, &B, gb1, &lb);
  for (int k = 0; k < 3; ++k) {
    g[k] = phi2 * ga1[k] * B + phi1 * phi2 * gb1[k];
    g[3 + k] = phi1 * ga2[k] * B - phi1 * phi2 * gb1[k];
  }
}

double orc_wf2_laplacian(const orc_wf2 *wf, const double cfg[6]) {
  double phi1, phi2, l1, l2, ga1[3], ga2[3], B, gb1[3], lb;
  orbital_vgl(wf, cfg, &phi1, ga1, &l1);
  orbital_vgl(wf, cfg + 3, &phi2, ga2, &l2);
  jastrow_vgl(wf, cfg, &B, gb1, &lb);
  double A = phi1 * phi2;
  /* sum_i (B lap_i A + A lap_i B + 2 grad_i A . grad_i B) */
  double dot1 = 0.0, dot2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    dot1 += phi2 * ga1[k] * gb1[k];
    dot2 += phi1 * ga2[k] * (-gb1[k]);
  }
  double lap = B * (phi2 * l1) + A * lb + 2.0 * dot1;
  lap += B * (phi1 * l2) + A * lb + 2.0 * dot2;
  return lap;
}

void orc_wf2_parameter_gradient(const orc_wf2 *wf, const double cfg[6],
                                double dpsi[ORC_NPARAM2]) {
  double zeta = wf->p[0], b = wf->p[1], c = wf->p[2];
  double psi = orc_wf2_value(wf, cfg);
  int nn = (wf->kind == ORC_WF_HELIUM_PADE) ? 1 : wf->n_nuc;
  /* d ln phi / d zeta = -sum_I r_I e_I / phi */
  double dz = 0.0;
  for (int e = 0; e < 2; ++e) {
    const double *x = cfg + 3 * e;
    double num = 0.0, phi = 0.0;
    for (int I = 0; I < nn; ++I) {
      double r = orc_norm3(x[0] - wf->nuc[I][0], x[1] - wf->nuc[I][1],
                           x[2] - wf->nuc[I][2]);
      double ee = orc_exp(-zeta * r);
      num += -r * ee;
      phi += ee;
    }
    dz += num / phi;
  }
  double r12 = orc_norm3(cfg[0] - cfg[3], cfg[1] - cfg[4], cfg[2] - cfg[5]);
  double den = 1.0 + c * r12;
  dpsi[0] = psi * dz;
  dpsi[1] = psi * (r12 / den);
  dpsi[2] = psi * (-b * r12 * r12 / (den * den));
}

double orc_ionic_potential(const orc_wf2 *wf, const double *cfg, int nelec) {
  double pot = 0.0;
  for (int i = 0; i < wf->n_nuc; ++i) {
    for (int j = 0; j < nelec; ++j) {
      double s = orc_norm3(cfg[3 * j] - wf->nuc[i][0],
                           cfg[3 * j + 1] - wf->nuc[i][1],
                           cfg[3 * j + 2] - wf->nuc[i][2]);
      pot -= (double)wf->charge[i] / s;
    }
  }
  /* proton-proton repulsion of IonicPotential::new (operator.rs:40-55) */
  double rep = 0.0;
  for (int i = 0; i < wf->n_nuc; ++i) {
    for (int j = i + 1; j < wf->n_nuc; ++j) {
      double s = orc_norm3(wf->nuc[j][0] - wf->nuc[i][0],
                           wf->nuc[j][1] - wf->nuc[i][1],
                           wf->nuc[j][2] - wf->nuc[i][2]);
      rep += (double)(wf->charge[i] * wf->charge[j]) / s;
    }
  }
  return pot + rep;
}

double orc_electronic_potential(const double *cfg, int nelec) {
  double pot = 0.0;
  for (int i = 0; i < nelec; ++i)
    for (int j = i + 1; j < nelec; ++j)
      pot += 1.0 / orc_norm3(cfg[3 * i] - cfg[3 * j],
                             cfg[3 * i + 1] - cfg[3 * j + 1],
                             cfg[3 * i + 2] - cfg[3 * j + 2]);
  return pot;
}

double orc_wf2_local_energy(const orc_wf2 *wf, const double cfg[6]) {
  /* act_on: KineticEnergy -0.5 lap (operator.rs:122-124), then
   * IonicPotential V psi (:58-61), then ElectronicPotential V psi (:93-96),
   * summed left to right (:181-182); Sampler::sample divides by psi. */
  double psi = orc_wf2_value(wf, cfg);
  double t = -0.5 * orc_wf2_laplacian(wf, cfg);
  double h = t + orc_ionic_potential(wf, cfg, 2) * psi;
  if (wf->hamiltonian == ORC_H_ELECTRONIC)
    h = h + orc_electronic_potential(cfg, 2) * psi;
  return h / psi;
}
