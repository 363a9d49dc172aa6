// Basis functions of the Slater-Jastrow orbitals about their centre R, at
// the displacement d = r - R with |d| = r (r from norm3 of d):
//   STO    exp(-z r)
//   GTO    exp(-z r^2)
//   STO_P  d_q exp(-z r), q = x, y, z (kinds MOLE_BASIS_STO_PX..PZ)
// with their gradients and laplacians
//   STO    grad = (-z e / r) d,                     lap = e (z^2 - 2 z / r)
//   GTO    grad = (-2 z e) d,                       lap = e (4 z^2 r^2 - 6 z)
//   STO_P  grad = ((-z e / r) d_q) d + e u_q,       lap = d_q e (z^2 - 4 z / r)
//          (each component adds e or +0)
// (u_q the unit vector; lap(x f(r)) = x (f'' + 4 f'/r)), with 1/r formed once
// and multiplied.  The oracle (sj.c)
// evaluates every form with the same operations in the same order, so
// values and gradients agree bit for bit.
#pragma once

#include "../../include/mole_gpu.h"
#include "device_common.hpp"

namespace mole {

__device__ __forceinline__ double pick_axis(int q, double d0, double d1, double d2) {
  return q == 0 ? d0 : (q == 1 ? d1 : d2);
}

// Value of a function of the given kind.  GTO: whether Gaussians can occur
// (false compiles their branch out).
template <bool GTO>
__device__ __forceinline__ double basis_value(int kind, double z, double d0, double d1, double d2, double r) {
  if (GTO && kind == MOLE_BASIS_GTO) return exp_m(-z * (r * r));
  const double e = exp_m(-z * r);
  if (kind >= MOLE_BASIS_STO_PX) return pick_axis(kind - MOLE_BASIS_STO_PX, d0, d1, d2) * e;
  return e;
}

// Value v, gradient (g0, g1, g2) and laplacian l.
template <bool GTO>
__device__ __forceinline__ void basis_vgl(int kind, double z, double d0, double d1, double d2, double r, double &v,
                                          double &g0, double &g1, double &g2, double &l) {
  if (GTO && kind == MOLE_BASIS_GTO) {
    const double r2 = r * r;
    v = exp_m(-z * r2);
    const double gf = -2.0 * z * v;
    g0 = gf * d0;
    g1 = gf * d1;
    g2 = gf * d2;
    l = v * (4.0 * z * z * r2 - 6.0 * z);
    return;
  }
  const double e = exp_m(-z * r);
  const double ir = 1.0 / r;  // one division
  const double gf = (-z * e) * ir;
  if (kind >= MOLE_BASIS_STO_PX) {
    const int q = kind - MOLE_BASIS_STO_PX;
    const double dq = pick_axis(q, d0, d1, d2);
    const double gq = gf * dq;
    v = dq * e;
    // + e on the function's own axis, + 0 on the others (selects, not a
    // dynamically indexed array)
    g0 = gq * d0 + (q == 0 ? e : 0.0);
    g1 = gq * d1 + (q == 1 ? e : 0.0);
    g2 = gq * d2 + (q == 2 ? e : 0.0);
    l = v * (z * z - (4.0 * z) * ir);
  } else {
    v = e;
    g0 = gf * d0;
    g1 = gf * d1;
    g2 = gf * d2;
    l = e * (z * z - (2.0 * z) * ir);
  }
}

// Gradient only (the drift): as basis_vgl's gradient.
template <bool GTO>
__device__ __forceinline__ void basis_grad(int kind, double z, double d0, double d1, double d2, double r, double &g0,
                                           double &g1, double &g2) {
  double v, l;
  basis_vgl<GTO>(kind, z, d0, d1, d2, r, v, g0, g1, g2, l);
}

// A determinant carried as mant 2^ex: the mantissa is renormalised into
// [1/2, 1) in magnitude by frexp after every product (exact), so products of
// many ratios neither overflow nor underflow.  log|det| = log|mant| + ex ln 2
// with ln 2 split as in log_m; a zero mantissa (a singular block) gives -inf.
__device__ __forceinline__ double logdet_value(double mant, double ex) {
  if (mant == 0.0) return -__builtin_inf();
  const double kLn2Hi = 6.93147180369123816490e-01;
  const double kLn2Lo = 1.90821492927058770002e-10;
  return fma(ex, kLn2Hi, fma(ex, kLn2Lo, log_m(fabs(mant))));
}

__device__ __forceinline__ void logdet_mul(double &mant, double &ex, double f) {
  // v_frexp_mant_f64 / v_frexp_exp_i32_f64: frexp's results for finite values
  const double x = mant * f;
  mant = __builtin_amdgcn_frexp_mant(x);
  ex = ex + (double)__builtin_amdgcn_frexp_exp(x);
}

}  // namespace mole
