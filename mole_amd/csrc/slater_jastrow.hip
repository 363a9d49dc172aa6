// Slater-Jastrow walker kernel: single-electron box moves with
// Sherman-Morrison determinant ratios and updates, periodic Gauss-Jordan
// re-inversion in FP64, and the local energy after sampled sweeps.
//
// A workgroup is one wave.  It owns 64/L walkers of one system (a tile);
// the L lanes of a walker split every stage of a move -- distances and
// Jastrow terms over electrons and nuclei, basis functions, orbitals,
// Sherman-Morrison columns and matrix elements -- and the walker's state
// (positions, inverse Slater matrices, scratch) lives in LDS for the whole
// launch, read from and written to HBM once.  Every sum is formed by the
// walker's sub-lane 0 in a fixed sequential order, so results do not depend
// on L and equal those of a sequential evaluation bit for bit.
//
// Math (per move of electron i of spin s, column c; B = A_s^-1 with
// A_s[c][k] = phi_k(r_c)):
//   ratio  R_D = sum_k phi_k(r') B[k][c]
//   dJ     = J_i(r') - J_i(r_i),  J_i = sum_{j != i} u_ij + sum_I v_I
//   accept min(R_D^2 exp(2 dJ), 1) > U           (metrop.rs:73-82)
//   update B[k][l] -= B[k][c] f_l (l != c), B[k][c] /= R_D,
//          f_l = (sum_k phi_k(r') B[k][l]) / R_D  (Sherman-Morrison)
#include "../../include/mole_gpu.h"
#include "basis.hpp"
#include "device_common.hpp"
#include "mlp.hpp"
#include "sj_internal.hpp"
#include "sr_math.hpp"

namespace mole {

// LDS layout of one walker (bytes); identical on host and device.
struct SJLayout {
  uint32_t pos, B, rng, rn, chi, phin, jn, jo, f, scr, flag, mv, acc, ld, bytes;
  __host__ __device__ SJLayout(int N, int nu, int nd, int nn, int K, int tsize, int np) {
    const int nmax = nu > nd ? nu : nd;
    uint32_t o = 0;
    auto take = [&](uint32_t n_bytes) {
      const uint32_t at = o;
      o += (n_bytes + 7u) & ~7u;
      return at;
    };
    pos = take(24u * N);
    B = take((uint32_t)tsize * (nu * nu + nd * nd));
    rng = take(32u * N);
    rn = take(8u * (nn > 0 ? nn : 1));
    chi = take(8u * K);
    phin = take(8u * nmax);
    jn = take(8u * (N + nn));
    jo = take(8u * (N + nn));
    f = take(8u * nmax);
    // scratch: Gauss-Jordan (2 n^2 + n), per-electron local-energy results
    // ((3 + np) N), drift pieces (3K + 3n + 3(N + n_nuc))
    const uint32_t lu = 8u * (2u * nmax * nmax + nmax);
    const uint32_t el = 8u * ((3u + (uint32_t)np) * N);
    const uint32_t dr = 8u * (3u * K + 3u * nmax + 3u * (N + nn));
    const uint32_t sc = lu > el ? lu : el;
    scr = take(sc > dr ? sc : dr);
    flag = take(16u);
    mv = take(64u);  // proposed position and drift (Gaussian proposals)
    acc = take(8u * sj_columns(np));
    ld = take(8u * kSJLdColumns);  // directly after acc: addressed as acc + sj_columns(np)
    // the walkers of a tile sit bytes apart; an odd number of 8-byte words
    // puts the same field of consecutive walkers in different LDS banks
    // (an even count maps many walkers of a wave onto one bank)
    bytes = ((o / 8u) % 2u == 0u) ? o + 8u : o;
  }
};

__host__ size_t sj_lds_per_walker(int N, int nu, int nd, int nn, int K, int tsize, int np) {
  return SJLayout(N, nu, nd, nn, K, tsize, np).bytes;
}

namespace {

template <typename T>
struct SJWalker {
  double *pos, *rng, *rn, *chi, *scr, *flag, *mv;
  T *Bup, *Bdn, *phin, *jn, *jo, *f;
};

__device__ __forceinline__ double pade_pair(double b, double c, double r) {
  return (b * r) / (1.0 + c * r);
}

// e-e term between electron i at r and electron j at x (T precision)
template <typename T>
__device__ __forceinline__ T pair_term(double b, double c, const double *r, const double *x);
template <>
__device__ __forceinline__ double pair_term<double>(double b, double c, const double *r, const double *x) {
  const double rij = norm3(r[0] - x[0], r[1] - x[1], r[2] - x[2]);
  return pade_pair(b, c, rij);
}
template <>
__device__ __forceinline__ float pair_term<float>(double b, double c, const double *r, const double *x) {
  const float d0 = (float)(r[0] - x[0]), d1 = (float)(r[1] - x[1]), d2 = (float)(r[2] - x[2]);
  float q = d0 * d0;
  q = q + d1 * d1;
  q = q + d2 * d2;
  const float rij = sqrtf(q);
  return ((float)b * rij) / (1.0f + (float)c * rij);
}

// e-n term of an electron at r and nucleus R with aZ = a Z_I (T precision)
template <typename T>
__device__ __forceinline__ T en_term(double aZ, double d, const double *r, const double *R);
template <>
__device__ __forceinline__ double en_term<double>(double aZ, double d, const double *r, const double *R) {
  const double rI = norm3(r[0] - R[0], r[1] - R[1], r[2] - R[2]);
  return -(aZ * rI) / (1.0 + d * rI);
}
template <>
__device__ __forceinline__ float en_term<float>(double aZ, double d, const double *r, const double *R) {
  const float d0 = (float)(r[0] - R[0]), d1 = (float)(r[1] - R[1]), d2 = (float)(r[2] - R[2]);
  float q = d0 * d0;
  q = q + d1 * d1;
  q = q + d2 * d2;
  const float rI = sqrtf(q);
  return -((float)aZ * rI) / (1.0f + (float)d * rI);
}

__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }

// e / n for 0 <= e < 4096, 1 <= n <= 64 via a float reciprocal: (e + 1/2)/n
// lies at least 1/(2n) from an integer, far beyond the rounding error.
__device__ __forceinline__ int small_div(int e, float inv_n) { return (int)(((float)e + 0.5f) * inv_n); }

// Basis values chi[m] of the point r into LDS (lanes over nuclei, then over
// basis functions).  gto/pk: Gaussians / p-type STOs may occur (otherwise
// every function is an s-type STO and the kinds are not read).  Ends with a
// barrier.
__device__ __forceinline__ void basis_to_lds(const SJSysDev &S, const double *r, double *rn, double *chi,
                                             int t, int L, bool gto, bool pk) {
  for (int I = t; I < S.n_nuc; I += L)
    rn[I] = norm3(r[0] - S.nuc[3 * I], r[1] - S.nuc[3 * I + 1], r[2] - S.nuc[3 * I + 2]);
  __syncthreads();
  for (int m = t; m < S.n_basis; m += L) {
    const int I = S.center[m];
    const double rr = rn[I];
    if (gto || pk) {
      const double *R = S.nuc + 3 * I;
      chi[m] = basis_value<true>(S.kind[m], S.zeta[m], r[0] - R[0], r[1] - R[1], r[2] - R[2], rr);
    } else {
      chi[m] = exp_m(-S.zeta[m] * rr);
    }
  }
  __syncthreads();
}

// Sequential sums whose loads are issued U at a time ahead of the
// arithmetic: the same operations in the same order as a plain loop, with
// the LDS/cache latency paid once per group instead of once per term.
template <int U, typename T>
__device__ __forceinline__ T seq_sum(const T *x, int n) {
  T s = (T)0;
  int j = 0;
  for (; j + U <= n; j += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = x[j + u];
#pragma unroll
    for (int u = 0; u < U; ++u) s = s + v[u];
  }
  for (; j < n; ++j) s = s + x[j];
  return s;
}

// sum_k a[k] b[k stride] as an fma chain in k order
template <int U, typename T>
__device__ __forceinline__ T seq_dot(const T *a, const T *b, int stride, int n) {
  T s = (T)0;
  int k = 0;
  for (; k + U <= n; k += U) {
    T av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      av[u] = a[k + u];
      bv[u] = b[(k + u) * stride];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s = fma_t(av[u], bv[u], s);
  }
  for (; k < n; ++k) s = fma_t(a[k], b[k * stride], s);
  return s;
}

// phi_k = sum_m c_km chi_m, fma chain in m order; coefficients from the
// [n_basis][n_orb] table (consecutive k on consecutive lanes)
__device__ __forceinline__ double orbital_value(const SJSysDev &S, int k, const double *chi) {
  const double *c = S.coef_t + k;
  const int K = S.n_basis, no = S.n_orb;
  double phi = 0.0;
  int m = 0;
  for (; m + 8 <= K; m += 8) {
    double cv[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      cv[u] = c[(size_t)(m + u) * no];
      xv[u] = chi[m + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) phi = fma(cv[u], xv[u], phi);
  }
  for (; m < K; ++m) phi = fma(c[(size_t)m * no], chi[m], phi);
  return phi;
}

// What a re-inversion of one spin block found (valid on sub-lane 0).
struct SJRefresh {
  double cond;      // ||A||_1 ||A^-1||_1
  double mant, ex;  // det A = mant 2^ex (mant = 0: singular)
  bool singular;
};

// Gauss-Jordan inverse of the spin-s Slater matrix from the positions, FP64,
// with partial pivoting, into B (as T) when write_B; the determinant from the
// pivots (the row swaps' signs times their product, carried as mant 2^ex).
// A pivot |p| <= n 2^-52 ||A||_1 (or NaN) marks the block singular: its
// division is skipped (the row is zeroed), so the inverse stays finite and
// the determinant reads 0.  All lanes of the workgroup must call it.
template <typename T>
__device__ SJRefresh refresh_spin(const SJSysDev &S, SJWalker<T> &wk, int s, int t, int L, bool gto, bool pk,
                                  bool write_B) {
  SJRefresh res{1.0, 1.0, 0.0, false};
  const int n = s == 0 ? S.n_up : S.n_dn;
  if (n == 0) return res;
  const int first = s == 0 ? 0 : S.n_up;
  T *B = s == 0 ? wk.Bup : wk.Bdn;
  double *M = wk.scr, *I = wk.scr + n * n, *colp = wk.scr + 2 * n * n;
  for (int c = 0; c < n; ++c) {
    basis_to_lds(S, wk.pos + 3 * (first + c), wk.rn, wk.chi, t, L, gto, pk);
    for (int k = t; k < n; k += L) M[c * n + k] = orbital_value(S, k, wk.chi);
    __syncthreads();
  }
  double anorm = 0.0;
  if (t == 0)
    for (int k = 0; k < n; ++k) {
      double cs = 0.0;
      for (int c = 0; c < n; ++c) cs += fabs(M[c * n + k]);
      anorm = cs > anorm ? cs : anorm;
    }
  const double thresh = anorm * ((double)n * 0x1p-52);
  const float inv_n = 1.0f / (float)n;
  for (int e = t; e < n * n; e += L) I[e] = (e == small_div(e, inv_n) * (n + 1)) ? 1.0 : 0.0;
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    if (t == 0) {
      int piv = p;
      for (int r = p + 1; r < n; ++r)
        if (fabs(M[r * n + p]) > fabs(M[piv * n + p])) piv = r;
      const double pv = M[piv * n + p];
      const bool bad = !(fabs(pv) > thresh);
      if (bad) {
        res.singular = true;
      } else {
        if (piv != p) res.mant = -res.mant;
        logdet_mul(res.mant, res.ex, pv);
      }
      wk.flag[0] = (double)piv;
      wk.flag[1] = bad ? 1.0 : 0.0;
    }
    __syncthreads();
    const int piv = (int)wk.flag[0];
    const bool bad = wk.flag[1] != 0.0;
    if (piv != p)
      for (int j = t; j < n; j += L) {
        double tm = M[p * n + j];
        M[p * n + j] = M[piv * n + j];
        M[piv * n + j] = tm;
        tm = I[p * n + j];
        I[p * n + j] = I[piv * n + j];
        I[piv * n + j] = tm;
      }
    __syncthreads();
    const double ip = bad ? 0.0 : 1.0 / M[p * n + p];
    __syncthreads();
    for (int j = t; j < n; j += L) {
      M[p * n + j] = M[p * n + j] * ip;
      I[p * n + j] = I[p * n + j] * ip;
    }
    for (int r = t; r < n; r += L) colp[r] = M[r * n + p];
    __syncthreads();
    // row p's own column entry was just scaled; colp[p] is unused
    for (int e = t; e < n * n; e += L) {
      const int r = small_div(e, inv_n), j = e - r * n;
      if (r == p) continue;
      M[e] = fma(-colp[r], M[p * n + j], M[e]);
      I[e] = fma(-colp[r], I[p * n + j], I[e]);
    }
    __syncthreads();
  }
  if (write_B)
    for (int e = t; e < n * n; e += L) B[e] = (T)I[e];
  double bnorm = 0.0;
  if (t == 0)
    for (int c = 0; c < n; ++c) {
      double cs = 0.0;
      for (int k = 0; k < n; ++k) cs += fabs(I[k * n + c]);
      bnorm = cs > bnorm ? cs : bnorm;
    }
  __syncthreads();
  res.cond = anorm * bnorm;
  if (res.singular) {
    res.mant = 0.0;
    res.ex = 0.0;
  }
  return res;
}

// Re-inverts both spin blocks and keeps the books on sub-lane 0: with count,
// the condition maximum and the singular blocks; the log-determinant state
// when ld_mode is 1 (taken from the re-inversion) or 2 (the same, after
// recording how far the value the updates carried drifted from it).
template <typename T>
__device__ void refresh_walker(const SJSysDev &S, SJWalker<T> &wk, double *acc, int t, int L, bool gto, bool pk,
                               bool write_B, bool count, int ld_mode) {
  double *ld = acc + sj_columns(5 + S.opt_exp);
#pragma unroll 1
  for (int q = 0; q < 2; ++q) {
    const SJRefresh r = refresh_spin<T>(S, wk, q, t, L, gto, pk, write_B);
    if (t == 0) {
      if (count) {
        acc[kSJCondMax] = fmax(acc[kSJCondMax], r.cond);
        if (r.singular) acc[kSJSingular] += 1.0;
      }
      if (ld_mode == 2 && !r.singular && ld[2 * q] != 0.0) {
        const double carried = logdet_value(ld[2 * q], ld[2 * q + 1]);
        acc[kSJLdDrift] = fmax(acc[kSJLdDrift], fabs(carried - logdet_value(r.mant, r.ex)));
      }
      if (ld_mode) {
        ld[2 * q] = r.mant;
        ld[2 * q + 1] = r.ex;
      }
    }
  }
}

// Local energy and d ln psi/dp (five Jastrow parameters) of the walker's
// configuration, whole electrons on lanes: lane t takes electrons t, t+L, ...
// and forms each electron's kinetic term (orbital gradients and laplacians
// as fma chains over the basis, contracted with the inverse column; Jastrow
// gradient and laplacian over partners, electrons then nuclei), potential and
// parameter derivatives in registers; sub-lane 0 adds the electrons' results
// in electron order.  Every sum has a fixed order, so the result does not
// depend on L.  One barrier per call; returns on sub-lane 0.
constexpr int kSJLanesNMax = 8;  // orbitals per register chunk

// The pieces of the local energy the operator terms need (sub-lane 0).
struct SJParts {
  double kin;  // sum_i lap_i psi / psi
  double vee;  // sum_{i<j} 1/r_ij
  double ven;  // -sum_{I,i} Z_I / |r_i - R_I|
};

// sum_k coef_k term_k of one operator (terms as the C ABI defines them).
__device__ double operator_value(const mole_operator &op, const SJParts &P, const SJSysDev &S, const double *pos) {
  const int N = S.N;
  double v = 0.0;
  for (uint32_t k = 0; k < op.n_terms; ++k) {
    const mole_term &tm = op.term[k];
    double x = 0.0;
    switch (tm.kind) {
      case MOLE_TERM_KINETIC: x = -0.5 * P.kin; break;
      case MOLE_TERM_IONIC_POTENTIAL: x = P.ven + S.v_nn; break;
      case MOLE_TERM_ELECTRONIC_POTENTIAL: x = P.vee; break;
      case MOLE_TERM_HARMONIC: {
        double s2 = 0.0;
        for (int i = 0; i < N; ++i)
          for (int q = 0; q < 3; ++q) {
            const double dq = pos[3 * i + q] - tm.param[1 + q];
            s2 += dq * dq;
          }
        x = 0.5 * (tm.param[0] * tm.param[0]) * s2;
        break;
      }
      case MOLE_TERM_RADIAL_POWER: {
        const int pw = (int)tm.param[0];
        for (int i = 0; i < N; ++i) {
          const double r = norm3(pos[3 * i] - tm.param[1], pos[3 * i + 1] - tm.param[2], pos[3 * i + 2] - tm.param[3]);
          const double b = pw > 0 ? r : 1.0 / r;
          double y = b;
          for (int e = 1; e < (pw > 0 ? pw : -pw); ++e) y *= b;
          x += y;
        }
        break;
      }
      case MOLE_TERM_DIPOLE:
        for (int i = 0; i < N; ++i)
          x += ((pos[3 * i] - tm.param[0]) * tm.param[3] + (pos[3 * i + 1] - tm.param[1]) * tm.param[4]) +
               (pos[3 * i + 2] - tm.param[2]) * tm.param[5];
        break;
      default: break;
    }
    v += tm.coef * x;
  }
  return v;
}

// Exponent scale s (zeta_m = s zeta0_m): sum_m zeta_m d ln psi / d zeta_m
// of electron i (at xi, column cc of its n x n inverse B), with d chi_m /
// d zeta_m = -r chi_m (s- and p-type STO) or -r^2 chi_m (GTO); fma chains in
// m, then k order.  Out of line: only systems with the scale call it, and the
// sweep kernels keep their code size.
template <typename T, bool EXT>
__device__ __noinline__ double exponent_scale_term(const SJSysDev &S, const T *B, int n, int cc, const double *xi) {
  const int K = S.n_basis, no = S.n_orb;
  double ez = 0.0;
  for (int k0 = 0; k0 < n; k0 += kSJLanesNMax) {
    double oz[kSJLanesNMax];
#pragma unroll
    for (int k = 0; k < kSJLanesNMax; ++k) oz[k] = 0.0;
    for (int m = 0; m < K; ++m) {
      const int I = S.center[m];
      const double d0 = xi[0] - S.nuc[3 * I], d1 = xi[1] - S.nuc[3 * I + 1], d2 = xi[2] - S.nuc[3 * I + 2];
      const double r = norm3(d0, d1, d2);
      const int kd = ((EXT && S.any_gto) || S.any_p) ? S.kind[m] : MOLE_BASIS_STO;
      const double ch = basis_value<EXT>(kd, S.zeta[m], d0, d1, d2, r);
      const double rp = (EXT && kd == MOLE_BASIS_GTO) ? r * r : r;
      const double tz = (-(S.zeta[m] * rp)) * ch;
      const double *cm = S.coef_t + (size_t)m * no + k0;
#pragma unroll
      for (int k = 0; k < kSJLanesNMax; ++k)
        if (k0 + k < n) oz[k] = fma(cm[k], tz, oz[k]);
    }
#pragma unroll
    for (int k = 0; k < kSJLanesNMax; ++k)
      if (k0 + k < n) ez = fma(oz[k], (double)B[(k0 + k) * n + cc], ez);
  }
  return ez;
}

// One compat sample of MonteCarloSampler::data (the extended
// instantiation only): "Wavefunction value" as Sampler::sample stores it,
// psi_s = (psi psi)/psi (src/vmc/src/operators.rs:19-23, samplers.rs:89-90),
// and "Parameter gradient" O_k psi_s with O = d ln psi/dp.  psi = sign
// exp(log|det up| + log|det dn| + J) from the carried determinants ld =
// (mant_up, ex_up, mant_dn, ex_dn) and J = b_anti O_0 + b_para O_1 + a O_3
// (the Pade factor is linear in b_anti, b_para and a).  Out of line: only
// contexts that keep compat samples call it.
__device__ __forceinline__ double sj_psi(const SJSysDev &S, const double *ld, const double *dl) {
  const double J = (S.p[0] * dl[0] + S.p[1] * dl[1]) + S.p[3] * dl[3];
  const double lg = (logdet_value(ld[0], ld[1]) + logdet_value(ld[2], ld[3])) + J;
  const double sg = ((ld[0] < 0.0) != (ld[2] < 0.0)) ? -1.0 : 1.0;
  return sg * exp_m(lg);
}

__device__ __noinline__ void sj_compat_sample(const SJSysDev &S, const double *ld, const double *dl, int np,
                                              size_t at) {
  const double psi = sj_psi(S, ld, dl);
  const double ps = (psi * psi) / psi;
  S.samp_psi[at] = ps;
  for (int k = 0; k < np; ++k) S.samp_g[at * (size_t)np + k] = dl[k] * ps;
}

// g_out (extended kernel, optional): grad_i ln psi of every electron, written
// by the lane that takes the electron as g_out[3 i + q].
template <typename T, bool EXT>
__device__ double local_energy(const SJSysDev &S, SJWalker<T> &wk, int t, int L, double *dl,
                               SJParts *parts = nullptr, double *g_out = nullptr) {
  const int N = S.N, K = S.n_basis, nn = S.n_nuc, no = S.n_orb;
  const double bA = S.p[0], bP = S.p[1], c = S.p[2], a = S.p[3], d = S.p[4];
  // [N][RS]: kinetic term, potential, five Jastrow derivatives, e-e potential
  // (EXT), exponent-scale derivative (with the scale); RS = 3 + np
  double *res = wk.scr;
  const int RS = 8 + S.opt_exp;
  SJParts own;
  for (int i = t; i < N; i += L) {
    const int s = i < S.n_up ? 0 : 1;
    const int cc = s == 0 ? i : i - S.n_up;
    const int n = s == 0 ? S.n_up : S.n_dn;
    const T *B = s == 0 ? wk.Bup : wk.Bdn;
    const double *xi = wk.pos + 3 * i;
    // orbitals in register chunks of kCh; the basis values are recomputed
    // per chunk (bit-identical each time), and the chunks add in orbital
    // order, so the width does not change the sums.  The extended
    // instantiations (DMC, operators, Gaussians) carry more live state and
    // take 6, which keeps their spills down (scratch 240 -> 176 bytes per
    // lane in the mixed-precision DMC step).
    constexpr int kCh = EXT ? 6 : kSJLanesNMax;
    double Dg0 = 0.0, Dg1 = 0.0, Dg2 = 0.0, Dl = 0.0;
    for (int k0 = 0; k0 < n; k0 += kCh) {
      double og[kCh][4];
#pragma unroll
      for (int k = 0; k < kCh; ++k) og[k][0] = og[k][1] = og[k][2] = og[k][3] = 0.0;
      for (int m = 0; m < K; ++m) {
        const int I = S.center[m];
        const double d0 = xi[0] - S.nuc[3 * I], d1 = xi[1] - S.nuc[3 * I + 1], d2 = xi[2] - S.nuc[3 * I + 2];
        const double r = norm3(d0, d1, d2);
        const int kd = ((EXT && S.any_gto) || S.any_p) ? S.kind[m] : MOLE_BASIS_STO;
        double ch, g0, g1, g2, lc;
        basis_vgl<EXT>(kd, S.zeta[m], d0, d1, d2, r, ch, g0, g1, g2, lc);
        const double *cm = S.coef_t + (size_t)m * no + k0;
#pragma unroll
        for (int k = 0; k < kCh; ++k)
          if (k0 + k < n) {
            const double cv = cm[k];
            og[k][0] = fma(cv, g0, og[k][0]);
            og[k][1] = fma(cv, g1, og[k][1]);
            og[k][2] = fma(cv, g2, og[k][2]);
            og[k][3] = fma(cv, lc, og[k][3]);
          }
      }
#pragma unroll
      for (int k = 0; k < kCh; ++k)
        if (k0 + k < n) {
          const double bkc = (double)B[(k0 + k) * n + cc];
          Dg0 = fma(og[k][0], bkc, Dg0);
          Dg1 = fma(og[k][1], bkc, Dg1);
          Dg2 = fma(og[k][2], bkc, Dg2);
          Dl = fma(og[k][3], bkc, Dl);
        }
    }
    // exponent scale s: sum_m zeta_m d ln psi / d zeta_m of electron i's row
    const double ez = S.opt_exp ? exponent_scale_term<T, EXT>(S, B, n, cc, xi) : 0.0;
    double gJ0 = 0.0, gJ1 = 0.0, gJ2 = 0.0, lJ = 0.0;
    double Vi = 0.0, Vee = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0;
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      const double *xj = wk.pos + 3 * j;
      const double dv0 = xi[0] - xj[0], dv1 = xi[1] - xj[1], dv2 = xi[2] - xj[2];
      const double r = norm3(dv0, dv1, dv2);
      const bool same = (j < S.n_up ? 0 : 1) == s;
      const double b = same ? bP : bA;
      // two divisions per pair: 1/r and 1/(1 + c r)
      const double ir = 1.0 / r;
      const double iden = 1.0 / (1.0 + c * r);
      const double iden2 = iden * iden;
      const double up = b * iden2;
      const double upp = (-2.0 * b * c) * (iden2 * iden);
      const double ur = up * ir;
      gJ0 += ur * dv0;
      gJ1 += ur * dv1;
      gJ2 += ur * dv2;
      lJ += upp + 2.0 * ur;
      if (j > i) {
        Vi += ir;
        if constexpr (EXT) Vee += ir;
        if (same)
          e1 += r * iden;
        else
          e0 += r * iden;
        e2 += (-b * r * r) * iden2;
      }
    }
    for (int I = 0; I < nn; ++I) {
      const double d0 = xi[0] - S.nuc[3 * I], d1 = xi[1] - S.nuc[3 * I + 1], d2 = xi[2] - S.nuc[3 * I + 2];
      const double r = norm3(d0, d1, d2);
      const double Z = S.charge[I];
      const double ir = 1.0 / r;
      const double iden = 1.0 / (1.0 + d * r);
      const double iden2 = iden * iden;
      const double aZ = a * Z;
      const double vp = -aZ * iden2;
      const double vpp = (2.0 * aZ * d) * (iden2 * iden);
      const double vr = vp * ir;
      gJ0 += vr * d0;
      gJ1 += vr * d1;
      gJ2 += vr * d2;
      lJ += vpp + 2.0 * vr;
      Vi += -Z * ir;
      e3 += (-Z * r) * iden;
      e4 += (aZ * r * r) * iden2;
    }
    const double dot = Dg0 * gJ0 + Dg1 * gJ1 + Dg2 * gJ2;
    const double gg = gJ0 * gJ0 + gJ1 * gJ1 + gJ2 * gJ2;
    if (EXT && g_out) {
      g_out[3 * i] = Dg0 + gJ0;
      g_out[3 * i + 1] = Dg1 + gJ1;
      g_out[3 * i + 2] = Dg2 + gJ2;
    }
    double *o = res + RS * i;
    o[0] = Dl + 2.0 * dot + lJ + gg;
    o[1] = Vi;
    o[2] = e0;
    o[3] = e1;
    o[4] = e2;
    o[5] = e3;
    o[6] = e4;
    if constexpr (EXT) o[7] = Vee;
    if (S.opt_exp) o[8] = ez;
  }
  __syncthreads();
  double kin = 0.0, V = 0.0;
  for (int q = 0; q < kSJNParam; ++q) dl[q] = 0.0;
  if (t == 0) {
    for (int i = 0; i < N; ++i) {
      const double *o = res + RS * i;
      kin += o[0];
      V += o[1];
      for (int q = 0; q < 5; ++q) dl[q] += o[2 + q];
      if (S.opt_exp) dl[5] += o[8];
    }
    // d ln psi / d s = (1/s) sum_m zeta_m d ln psi / d zeta_m
    if (S.opt_exp) dl[5] = dl[5] / S.p[5];
    if (S.hamiltonian != MOLE_H_ELECTRONIC) {
      V = 0.0;
      for (int i = 0; i < N; ++i)
        for (int I = 0; I < nn; ++I) {
          const double *xi = wk.pos + 3 * i;
          V += -S.charge[I] / norm3(xi[0] - S.nuc[3 * I], xi[1] - S.nuc[3 * I + 1], xi[2] - S.nuc[3 * I + 2]);
        }
    }
    if (EXT && parts) parts->kin = kin;
    if (EXT && (S.custom_h || S.n_obs)) {
      if (!parts) parts = &own;
      parts->kin = kin;
      double vee = 0.0, ven = 0.0;
      for (int i = 0; i < N; ++i) vee += res[RS * i + 7];
      // IonicPotential::value's order: ions outer, electrons inner
      for (int I = 0; I < nn; ++I)
        for (int i = 0; i < N; ++i) {
          const double *xi = wk.pos + 3 * i;
          ven -= S.charge[I] / norm3(xi[0] - S.nuc[3 * I], xi[1] - S.nuc[3 * I + 1], xi[2] - S.nuc[3 * I + 2]);
        }
      parts->vee = vee;
      parts->ven = ven;
    }
  }
  __syncthreads();
  if (EXT && t == 0 && S.custom_h) return operator_value(S.ops[0], *parts, S, wk.pos);
  return -0.5 * kin + V + S.v_nn;
}

// Pieces of grad_i ln psi at r for the drift-diffusion proposal, FP64
// (metrop.rs:153-155).  Basis gradients into gch[3m + q] (lanes over basis
// functions; chi and rn of r must be in LDS) and the Jastrow gradient terms
// of electron i at r into jg[3j + q] (lanes over partners, i's slot zero).
// No barrier.
__device__ __forceinline__ void drift_terms(const SJSysDev &S, const double *pos, int i, const double *r,
                                            const double *rn, const double *chi, double *gch, double *jg,
                                            int t, int L, bool gto, bool pk) {
  for (int m = t; m < S.n_basis; m += L) {
    const int I = S.center[m];
    const double d0 = r[0] - S.nuc[3 * I], d1 = r[1] - S.nuc[3 * I + 1], d2 = r[2] - S.nuc[3 * I + 2];
    const int kd = (gto || pk) ? S.kind[m] : MOLE_BASIS_STO;
    if (kd == MOLE_BASIS_STO) {
      // chi[m] = exp(-z r) is in LDS; basis_vgl's operations
      const double gf = (-S.zeta[m] * chi[m]) * (1.0 / rn[I]);
      gch[3 * m] = gf * d0;
      gch[3 * m + 1] = gf * d1;
      gch[3 * m + 2] = gf * d2;
    } else {
      basis_grad<true>(kd, S.zeta[m], d0, d1, d2, rn[I], gch[3 * m], gch[3 * m + 1], gch[3 * m + 2]);
    }
  }
  const int N = S.N, si = i < S.n_up ? 0 : 1;
  for (int j = t; j < N + S.n_nuc; j += L) {
    double *o = jg + 3 * j;
    if (j < N) {
      if (j == i) {
        o[0] = o[1] = o[2] = 0.0;
        continue;
      }
      const double *xj = pos + 3 * j;
      const double dv0 = r[0] - xj[0], dv1 = r[1] - xj[1], dv2 = r[2] - xj[2];
      const double rr = norm3(dv0, dv1, dv2);
      const double b = ((j < S.n_up ? 0 : 1) == si) ? S.p[1] : S.p[0];
      const double den = 1.0 + S.p[2] * rr;
      const double up = b / (den * den);
      o[0] = up * dv0 / rr;
      o[1] = up * dv1 / rr;
      o[2] = up * dv2 / rr;
    } else {
      const int I = j - N;
      const double d0 = r[0] - S.nuc[3 * I], d1 = r[1] - S.nuc[3 * I + 1], d2 = r[2] - S.nuc[3 * I + 2];
      const double rr = norm3(d0, d1, d2);
      const double den = 1.0 + S.p[4] * rr;
      const double vp = -(S.p[3] * S.charge[I]) / (den * den);
      o[0] = vp * d0 / rr;
      o[1] = vp * d1 / rr;
      o[2] = vp * d2 / rr;
    }
  }
}

// Orbital gradients og[3k + q] = sum_m c_km gch[3m + q] (lanes over
// orbitals, chains in m order).  No barrier.
__device__ __forceinline__ void orbital_gradients(const SJSysDev &S, int n, const double *gch, double *og,
                                                  int t, int L) {
  const int K = S.n_basis, no = S.n_orb;
  for (int k = t; k < n; k += L) {
    const double *cf = S.coef_t + k;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int m = 0; m < K; ++m) {
      const double cm = cf[(size_t)m * no];
      g0 = fma(cm, gch[3 * m], g0);
      g1 = fma(cm, gch[3 * m + 1], g1);
      g2 = fma(cm, gch[3 * m + 2], g2);
    }
    og[3 * k] = g0;
    og[3 * k + 1] = g1;
    og[3 * k + 2] = g2;
  }
}

// Sub-lane 0: g = sum_k og_k B[k][c] (chains in k order) + sum_j jg_j (in j
// order), each as in the CPU restatement.
template <typename T>
__device__ __forceinline__ void drift_sum(const double *og, const T *B, int n, int c, const double *jg, int M,
                                          double det_scale, double g[3]) {
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int k = 0; k < n; ++k) {
    const double bkc = (double)B[k * n + c];
    d0 = fma(og[3 * k], bkc, d0);
    d1 = fma(og[3 * k + 1], bkc, d1);
    d2 = fma(og[3 * k + 2], bkc, d2);
  }
  double j0 = 0.0, j1 = 0.0, j2 = 0.0;
  for (int j = 0; j < M; ++j) {
    j0 += jg[3 * j];
    j1 += jg[3 * j + 1];
    j2 += jg[3 * j + 2];
  }
  // det_scale = 1: plain sum; otherwise the determinant part is divided by it
  if (det_scale != 1.0) {
    d0 = d0 / det_scale;
    d1 = d1 / det_scale;
    d2 = d2 / det_scale;
  }
  g[0] = d0 + j0;
  g[1] = d1 + j1;
  g[2] = d2 + j2;
}

}  // namespace

// EXT: the DMC modes, local operators and Gaussian basis functions are
// compiled in (the plain VMC instantiations leave them out, keeping their
// register budget).
// MINW: minimum waves per SIMD the register budget is fitted to (2: up to
// 256 registers; 1: 512, for the learned proposal's larger systems, where
// spills cost more than the lost occupancy).
// One tile's launch: its walkers' state into the workgroup's LDS, the sweeps,
// the state out.
template <typename T, int PROP, bool EXT>
__device__ __forceinline__ void sj_tile(const SJArgs &a, const SJTile tile, unsigned char *lds) {
  constexpr int DPM = sj_draws_per_move(PROP);
  const SJSysDev &S = a.systems[tile.sys];
  const int N = S.N, nu = S.n_up, nd = S.n_dn, L = S.L, W = S.W;
  const int lane = threadIdx.x;
  // learned proposal: 16 walkers of 4 lanes, walker s on lanes s, s+16,
  // s+32, s+48 (the sample columns of the MFMA tiles); otherwise L
  // consecutive lanes per walker
  const int slot = PROP == kSJPropLearned ? (lane & 15) : lane / L;
  const int t = PROP == kSJPropLearned ? (lane >> 4) : lane % L;
  const int nslots = 64 / L;
  const int lg_slots = __ffs(nslots) - 1;  // nslots is a power of two
  const bool live = slot < tile.nw;
  const int w = tile.w0 + slot;
  const int np = 5 + S.opt_exp, ncol = sj_columns(np);
  const SJLayout lay(N, nu, nd, S.n_nuc, S.n_basis, (int)sizeof(T), np);
  unsigned char *base = lds + (size_t)slot * lay.bytes;
  SJWalker<T> wk;
  wk.pos = (double *)(base + lay.pos);
  wk.Bup = (T *)(base + lay.B);
  wk.Bdn = wk.Bup + nu * nu;
  wk.rng = (double *)(base + lay.rng);
  wk.rn = (double *)(base + lay.rn);
  wk.chi = (double *)(base + lay.chi);
  wk.phin = (T *)(base + lay.phin);
  wk.jn = (T *)(base + lay.jn);
  wk.jo = (T *)(base + lay.jo);
  wk.f = (T *)(base + lay.f);
  wk.scr = (double *)(base + lay.scr);
  wk.flag = (double *)(base + lay.flag);
  wk.mv = (double *)(base + lay.mv);
  const int nB = nu * nu + nd * nd;
  const bool gto = EXT && S.any_gto;
  const bool pk = S.any_p != 0;
  // DMC steps move walkers the branching gathered: the log-determinant state
  // is not carried there (mole_dmc_diffuse recomputes it at its end)
  const bool track_ld = !(EXT && a.dmc == kSJDmcStep);

  // ---- state in (a DMC step reads the walkers the branching gathered) ----
  const bool dmc_in = EXT && a.dmc == kSJDmcStep;
  const double *pin = dmc_in ? S.dmc_pos_in : S.pos;
  const T *Bg = (const T *)(dmc_in ? S.dmc_B_in : S.B);
  for (int e = lane; e < 3 * N * nslots; e += 64) {
    const int k = e >> lg_slots, sl = e & (nslots - 1);
    double *pp = (double *)(lds + (size_t)sl * lay.bytes + lay.pos);
    pp[k] = (sl < tile.nw) ? pin[(size_t)k * W + tile.w0 + sl] : 0.0;
  }
  for (int e = lane; e < nB * nslots; e += 64) {
    const int k = e >> lg_slots, sl = e & (nslots - 1);
    T *bb = (T *)(lds + (size_t)sl * lay.bytes + lay.B);
    bb[k] = (sl < tile.nw) ? Bg[(size_t)k * W + tile.w0 + sl] : (T)0;
  }
  // accumulators: LDS, touched by sub-lane 0 only
  double *acc = (double *)(base + lay.acc);
  if (t == 0) {
    // DMC steps touch only the accepted-move count and the condition maximum
    const bool dmc_step = EXT && a.dmc == kSJDmcStep;
    for (int k = 0; k < ncol; ++k)
      if (!dmc_step || k == kSJAccepted || k == kSJCondMax || k == kSJSingular)
        acc[k] = live ? S.acc[(size_t)k * W + w] : 0.0;
    if (track_ld)
      for (int k = 0; k < kSJLdColumns; ++k) acc[ncol + k] = live ? S.ld[(size_t)k * W + w] : 1.0;
  }
  __syncthreads();

  const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32) ^ S.sys_index;
  if (a.init) {
    // Sampler::new: U(-1, 1) from draws 0..3N-1 (samplers.rs:46)
    for (int b = t; 2 * b < 3 * N; b += L) {
      const u32x4 o = philox4x32_10({(uint32_t)b, 0u, (uint32_t)w, a.epoch}, k0, k1);
      const uint64_t v0 = (uint64_t)o.x | ((uint64_t)o.y << 32);
      const uint64_t v1 = (uint64_t)o.z | ((uint64_t)o.w << 32);
      wk.pos[2 * b] = u64_range(v0, 2.0, -3.0);
      if (2 * b + 1 < 3 * N) wk.pos[2 * b + 1] = u64_range(v1, 2.0, -3.0);
    }
    __syncthreads();
  }
  if (a.init || a.refresh) {
    // new configurations: the determinants are taken from the re-inversion
    // (nothing accumulated to compare with); refresh 2 keeps the inverses
    const bool wb = a.init || a.refresh == 1;
    refresh_walker<T>(S, wk, acc, t, L, gto, pk, wb, wb, track_ld ? 1 : 0);
  }
  if (EXT && a.dmc == kSJDmcInit) {
    // DmcRunner::new: weight 1; the local energy of the starting configuration
    double dlg[kSJNParam];
    const double E = local_energy<T, EXT>(S, wk, t, L, dlg);
    if (t == 0 && live) {
      S.dmc[w] = 1.0;
      S.dmc[(size_t)W + w] = E;
    }
  }
  if (EXT && a.dmc == kSJEval) {
    // mole_eval: psi = sign exp(log|det up| + log|det dn| + J) from the
    // re-inversion's determinants, d psi = psi grad ln psi, lap psi = psi
    // sum_i lap_i psi/psi, E_L and d psi/dp = psi d ln psi/dp
    double dlg[kSJNParam];
    SJParts parts;
    double *gw = (live && S.ev_grad) ? S.ev_grad + (size_t)w * 3 * N : nullptr;
    const double E = local_energy<T, EXT>(S, wk, t, L, dlg, &parts, gw);
    if (t == 0 && live) {
      const double psi = sj_psi(S, acc + ncol, dlg);
      if (S.ev_value) S.ev_value[w] = psi;
      if (S.ev_lap) S.ev_lap[w] = psi * parts.kin;
      if (S.ev_el) S.ev_el[w] = E;
      if (S.ev_pgrad)
        for (int k = 0; k < np; ++k) S.ev_pgrad[(size_t)w * np + k] = psi * dlg[k];
      // the lanes' gradient stores precede local_energy's closing barrier
      if (gw)
        for (int k = 0; k < 3 * N; ++k) gw[k] = psi * gw[k];
    }
  }
  const uint32_t lu_period = a.lu_period;

  const SJLaunchState ls = *a.state;
  uint64_t p = S.draw_base + DPM * (uint64_t)N * ls.sweeps_done;
  uint64_t m = S.move_base + (uint64_t)N * ls.sweeps_done;
  uint32_t fill = ls.block_fill;
  uint32_t sidx = ls.trace_samples;
  const double inv_block = 1.0 / (double)a.block_size;

  for (uint32_t sw = 0; sw < a.n_sweeps; ++sw) {
    if constexpr (PROP == kSJPropBox) {
      // ---- the sweep's draws: [p, p + 4N), lanes over Philox blocks ----
      const uint64_t b0 = p >> 1, b1 = (p + 4 * (uint64_t)N - 1) >> 1;
      for (uint64_t b = b0 + t; b <= b1; b += L) {
        const u32x4 o = philox4x32_10({(uint32_t)b, (uint32_t)(b >> 32), (uint32_t)w, a.epoch}, k0, k1);
        const uint64_t v[2] = {(uint64_t)o.x | ((uint64_t)o.y << 32), (uint64_t)o.z | ((uint64_t)o.w << 32)};
        for (int h = 0; h < 2; ++h) {
          const uint64_t q = 2 * b + h;
          if (q < p || q >= p + 4 * (uint64_t)N) continue;
          const int r = (int)(q - p);
          wk.rng[r] = ((r & 3) == 3) ? u64_standard(v[h]) : u64_range(v[h], a.scale, a.offset);
        }
      }
    } else {
      // ---- Gaussian proposals: 6 draws per move, lanes over moves; Box-Muller
      // of draws (0,1) and (2,3) -> z0, z1, z2; draw 4 -> the accept uniform
      Stream st;
      st.init(a.seed, S.sys_index, (uint32_t)w, a.epoch);
      for (int i = t; i < N; i += L) {
        const uint64_t q = p + 6 * (uint64_t)i;
        const uint64_t d0 = st.draw(q), d1 = st.draw(q + 1), d2 = st.draw(q + 2), d3 = st.draw(q + 3);
        const uint64_t d4 = st.draw(q + 4);
        double z0, z1, z2, z3;
        box_muller(d0, d1, &z0, &z1);
        box_muller(d2, d3, &z2, &z3);
        wk.rng[4 * i] = z0;
        wk.rng[4 * i + 1] = z1;
        wk.rng[4 * i + 2] = z2;
        wk.rng[4 * i + 3] = u64_standard(d4);
      }
    }
    __syncthreads();
    uint64_t mask = 0;
    for (int i = 0; i < N; ++i) {
      const int s = i < nu ? 0 : 1;
      const int c = s == 0 ? i : i - nu;
      const int n = s == 0 ? nu : nd;
      T *B = s == 0 ? wk.Bup : wk.Bdn;
      const double *xi = wk.pos + 3 * i;
      double xn[3];
      double *gch = wk.scr, *og = wk.scr + 3 * S.n_basis, *jg = og + 3 * (nu > nd ? nu : nd);
      if constexpr (PROP == kSJPropBox) {
        xn[0] = xi[0] + wk.rng[4 * i];
        xn[1] = xi[1] + wk.rng[4 * i + 1];
        xn[2] = xi[2] + wk.rng[4 * i + 2];
      } else if constexpr (PROP == kSJPropLearned) {
        // learned proposal: pass 0, the network at x -> (mu, ln sigma), x' =
        // (x_i + mu) + sigma z; pass 1, the network at x' for the reverse
        // density (all lanes run the MFMA passes).  One loop body keeps a
        // single copy of the features and the network in the kernel's code.
#pragma unroll 1
        for (int pass = 0; pass < 2; ++pass) {
          float fv[4];
          mlp_features(wk.pos, N, S.nuc, S.n_nuc, i, pass == 0 ? xi : wk.mv, t, fv);
          const f32x4 o = mlp_forward(a.net, features_bf16(fv), nullptr);
          if (t == 0) {
            const double ls = mlp_log_sigma((double)o[3]);
            const double sg = exp_m(ls);
            if (pass == 0) {
              double z2 = 0.0;
              for (int q = 0; q < 3; ++q) {
                const double z = wk.rng[4 * i + q];
                wk.mv[q] = (xi[q] + (double)o[q]) + sg * z;
                z2 += z * z;
              }
              wk.mv[3] = -3.0 * ls - 0.5 * z2;  // log q(x'|x) + const
            } else {
              double e2 = 0.0;
              for (int q = 0; q < 3; ++q) {
                const double e = (xi[q] - wk.mv[q]) - (double)o[q];
                e2 += e * e;
              }
              wk.mv[4] = (-3.0 * ls - e2 / (2.0 * sg * sg)) - wk.mv[3];  // log q(x|x') - log q(x'|x)
            }
          }
          __syncthreads();
        }
        xn[0] = wk.mv[0];
        xn[1] = wk.mv[1];
        xn[2] = wk.mv[2];
      } else {
        // drift at x (metrop.rs:150-163): x' = (x_i + tau v) + sqrt(tau) z
        basis_to_lds(S, xi, wk.rn, wk.chi, t, L, gto, pk);
        drift_terms(S, wk.pos, i, xi, wk.rn, wk.chi, gch, jg, t, L, gto, pk);
        __syncthreads();
        orbital_gradients(S, n, gch, og, t, L);
        __syncthreads();
        if (t == 0) {
          double v[3];
          drift_sum<T>(og, B, n, c, jg, N + S.n_nuc, 1.0, v);
          for (int q = 0; q < 3; ++q) {
            wk.mv[3 + q] = v[q];
            wk.mv[q] = (xi[q] + v[q] * a.tau) + a.sqrt_tau * wk.rng[4 * i + q];
          }
        }
        __syncthreads();
        xn[0] = wk.mv[0];
        xn[1] = wk.mv[1];
        xn[2] = wk.mv[2];
      }
      // Jastrow terms of electron i at r' and at r_i (lanes over partners)
      for (int j = t; j < N + S.n_nuc; j += L) {
        if (j < N) {
          if (j == i) {  // zero: the sums below start at +0, so it adds nothing
            wk.jn[j] = (T)0;
            wk.jo[j] = (T)0;
            continue;
          }
          const double b = ((j < nu ? 0 : 1) == s) ? S.p[1] : S.p[0];
          wk.jn[j] = pair_term<T>(b, S.p[2], xn, wk.pos + 3 * j);
          wk.jo[j] = pair_term<T>(b, S.p[2], xi, wk.pos + 3 * j);
        } else {
          const int I = j - N;
          const double aZ = S.p[3] * S.charge[I];
          wk.jn[j] = en_term<T>(aZ, S.p[4], xn, S.nuc + 3 * I);
          wk.jo[j] = en_term<T>(aZ, S.p[4], xi, S.nuc + 3 * I);
        }
      }
      basis_to_lds(S, xn, wk.rn, wk.chi, t, L, gto, pk);
      for (int k = t; k < n; k += L) wk.phin[k] = (T)orbital_value(S, k, wk.chi);
      if constexpr (PROP == kSJPropDiffuse) {
        drift_terms(S, wk.pos, i, xn, wk.rn, wk.chi, gch, jg, t, L, gto, pk);
        __syncthreads();
        orbital_gradients(S, n, gch, og, t, L);
      }
      __syncthreads();
      if (t == 0) {
        const T rd = seq_dot<8>(wk.phin, B + c, n, n);
        const T Jn = seq_sum<8>(wk.jn, N + S.n_nuc);
        const T Jo = seq_sum<8>(wk.jo, N + S.n_nuc);
        const T dJ = Jn - Jo;
        const double rdd = (double)rd;
        const double R2 = (rdd * rdd) * exp_m(2.0 * (double)dJ);
        bool accept;
        if constexpr (PROP == kSJPropBox) {
          accept = min_one(R2) > wk.rng[4 * i + 3];
        } else if constexpr (PROP == kSJPropLearned) {
          // exact MH ratio of the learned proposal
          accept = min_one(R2 * exp_m(wk.mv[4])) > wk.rng[4 * i + 3];
        } else {
          // metrop.rs:178-197 over electron i's row: reject a sign change,
          // else min(T(x'->x) R^2 / T(x->x'), 1) > u
          double vn[3];
          drift_sum<T>(og, B, n, c, jg, N + S.n_nuc, rdd, vn);
          double eh[3], el[3];
          for (int q = 0; q < 3; ++q) {
            eh[q] = (xi[q] - xn[q]) - vn[q] * a.tau;
            el[q] = (xn[q] - xi[q]) - wk.mv[3 + q] * a.tau;
          }
          const double nh = norm3(eh[0], eh[1], eh[2]);
          const double nl = norm3(el[0], el[1], el[2]);
          const double t_high = exp_m(-(nh * nh) / (2.0 * a.tau));
          const double t_low = exp_m(-(nl * nl) / (2.0 * a.tau));
          accept = !(rdd < 0.0) && min_one((t_high * R2) / t_low) > wk.rng[4 * i + 3];
        }
        wk.flag[0] = accept ? 1.0 : 0.0;
        wk.flag[1] = (double)((T)1 / rd);
        // det A_s' = R_D det A_s, carried in FP64
        if (accept && track_ld) logdet_mul(acc[ncol + 2 * s], acc[ncol + 2 * s + 1], rdd);
      }
      __syncthreads();
      const bool accept = wk.flag[0] != 0.0;
      const T inv = (T)wk.flag[1];
      if (accept)
        for (int l = t; l < n; l += L) {
          if (l == c) continue;
          wk.f[l] = seq_dot<8>(wk.phin, B + l, n, n) * inv;
        }
      __syncthreads();
      if (accept) {
        // B[k][l] -= B[k][c] f_l for l != c; loads of a group ahead of its
        // stores (column c is only read here)
        const float inv_n = 1.0f / (float)n;
        for (int e0 = t; e0 < n * n; e0 += 4 * L) {
          T bkc[4], fl[4], be[4];
          int ee[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * L;
            const int k = small_div(e, inv_n), l = e - k * n;
            ee[u] = (e < n * n && l != c) ? e : -1;
            if (ee[u] >= 0) {
              bkc[u] = B[k * n + c];
              fl[u] = wk.f[l];
              be[u] = B[e];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (ee[u] >= 0) B[ee[u]] = fma_t(-bkc[u], fl[u], be[u]);
        }
      }
      __syncthreads();
      if (accept) {
        for (int k = t; k < n; k += L) B[k * n + c] = B[k * n + c] * inv;
        if (t == 0) {
          wk.pos[3 * i] = xn[0];
          wk.pos[3 * i + 1] = xn[1];
          wk.pos[3 * i + 2] = xn[2];
          mask |= (uint64_t)1 << i;
          acc[kSJAccepted] += 1.0;
        }
      }
      __syncthreads();
      ++m;
      if (lu_period && m % lu_period == 0) {
        // the FP64 log|det| the update chain carried is checked against the re-inversion's
        refresh_walker<T>(S, wk, acc, t, L, gto, pk, true, true, track_ld ? 2 : 0);
      }
    }
    p += DPM * (uint64_t)N;
    if (t == 0 && live && S.trace_mask && ls.trace_sweeps + sw < a.cap_sweeps)
      S.trace_mask[(size_t)w * a.cap_sweeps + ls.trace_sweeps + sw] = mask;
    if (EXT && sw >= a.sample_start && a.dmc == kSJDmcStep) {
      // DmcRunner::diffuse (src/dmc/src/dmc.rs:88-128): the ensemble terms
      // with the weight before the step, then
      // w *= exp(-tau ((E_old + E_new)/2 - E_ref))
      double dlg[kSJNParam];
      const double E = local_energy<T, EXT>(S, wk, t, L, dlg);
      if (t == 0 && live) {
        const double eo = S.dmc_el_in[w], wt = S.dmc_w_in[w];
        S.dmc[2 * (size_t)W + w] = wt * eo;
        S.dmc[3 * (size_t)W + w] = wt;
        S.dmc[w] = wt * exp_m(-a.dmc_tau * ((eo + E) / 2.0 - *S.dmc_eref));
        S.dmc[(size_t)W + w] = E;
      }
    } else if (sw >= a.sample_start) {
      double dlg[kSJNParam];
      SJParts parts;
      const double E = local_energy<T, EXT>(S, wk, t, L, dlg, &parts);
      if (EXT && t == 0 && live && S.n_obs) {
        // further observables: per-walker sums in HBM (opt-in, sampled sweeps only)
        for (int k = 0; k < S.n_obs; ++k) {
          const double v = operator_value(S.ops[1 + k], parts, S, wk.pos);
          double *o = S.obs_acc + (size_t)(2 * k) * W + w;
          o[0] += v;
          o[(size_t)W] = fma(v, v, o[(size_t)W]);
        }
      }
      if (t == 0) {
        acc[kSJSumE] += E;
        acc[kSJSumE2] = fma(E, E, acc[kSJSumE2]);
        acc[kSJBlockSum] += E;
        int q = 0;
        for (int k = 0; k < np; ++k) {
          acc[kSJSumO + k] += dlg[k];
          acc[sj_sum_oe(np) + k] = fma(dlg[k], E, acc[sj_sum_oe(np) + k]);
          for (int l = k; l < np; ++l, ++q) acc[sj_sum_oo(np) + q] = fma(dlg[k], dlg[l], acc[sj_sum_oo(np) + q]);
        }
        if (live && S.trace_e && sidx < a.cap_samples) S.trace_e[(size_t)w * a.cap_samples + sidx] = E;
        if (EXT && live && S.samp_psi && sidx < a.cap_samples)
          sj_compat_sample(S, acc + ncol, dlg, np, (size_t)w * a.cap_samples + sidx);
      }
      if (++fill == a.block_size) {
        if (t == 0) {
          const double bm = acc[kSJBlockSum] * inv_block;
          acc[kSJSumBm] += bm;
          acc[kSJSumBm2] = fma(bm, bm, acc[kSJSumBm2]);
          acc[kSJBlockSum] = 0.0;
        }
        fill = 0;
      }
      ++sidx;
    }
  }

  // ---- state out ----
  __syncthreads();
  for (int e = lane; e < 3 * N * nslots; e += 64) {
    const int k = e >> lg_slots, sl = e & (nslots - 1);
    if (sl < tile.nw) S.pos[(size_t)k * W + tile.w0 + sl] = ((double *)(lds + (size_t)sl * lay.bytes + lay.pos))[k];
  }
  T *Bo = (T *)S.B;
  for (int e = lane; e < nB * nslots; e += 64) {
    const int k = e >> lg_slots, sl = e & (nslots - 1);
    if (sl < tile.nw) Bo[(size_t)k * W + tile.w0 + sl] = ((T *)(lds + (size_t)sl * lay.bytes + lay.B))[k];
  }
  if (t == 0 && live) {
    for (int k = 0; k < ncol; ++k)
      if (!(EXT && a.dmc == kSJDmcStep) || k == kSJAccepted || k == kSJCondMax || k == kSJSingular)
        S.acc[(size_t)k * W + w] = acc[k];
    if (track_ld)
      for (int k = 0; k < kSJLdColumns; ++k) S.ld[(size_t)k * W + w] = acc[ncol + k];
  }
}

// One workgroup per tile (the tiles of one LDS footprint class).
template <typename T, int PROP, bool EXT, int MINW>
__global__ void __launch_bounds__(64, MINW) sj_sweep_kernel(SJArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  sj_tile<T, PROP, EXT>(a, a.tiles[blockIdx.x], lds);
}

// Persistent: a grid sized to the resident workgroups takes every tile of the
// context, longest first, each workgroup drawing the next tile index from a
// device ticket until the list is exhausted (every wave reaches the exit).
// The ticket is zero on entry; sj_ticket_reset_kernel zeroes it after.
template <typename T>
__global__ void __launch_bounds__(64, 2) sj_persistent_kernel(SJArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (;;) {
    int t = 0;
    if (threadIdx.x == 0) t = (int)atomicAdd(a.ticket, 1u);
    t = __shfl(t, 0);
    if (t >= a.n_tiles) break;
    sj_tile<T, kSJPropBox, false>(a, a.tiles[t], lds);
    __syncthreads();  // the next tile reuses the LDS
  }
}

__global__ void sj_ticket_reset_kernel(uint32_t *ticket) {
  if (threadIdx.x == 0) *ticket = 0u;
}

__global__ void sj_advance_kernel(SJLaunchState *st, uint32_t n_sweeps, uint32_t n_samples,
                                  uint32_t block_size) {
  if (threadIdx.x != 0) return;
  st->sweeps_done += n_sweeps;
  st->block_fill = (uint32_t)(((uint64_t)st->block_fill + n_samples) % block_size);
  st->trace_sweeps += n_sweeps;
  st->trace_samples += n_samples;
}

__global__ void sj_state_set_kernel(SJLaunchState *st, SJLaunchState v) {
  if (threadIdx.x == 0) *st = v;
}

hipError_t launch_sj_state_set(SJLaunchState *state, const SJLaunchState &v, hipStream_t stream) {
  hipLaunchKernelGGL(sj_state_set_kernel, dim3(1), dim3(64), 0, stream, state, v);
  return hipGetLastError();
}

hipError_t launch_sj_advance(SJLaunchState *state, uint32_t n_sweeps, uint32_t n_samples,
                             uint32_t block_size, hipStream_t stream) {
  hipLaunchKernelGGL(sj_advance_kernel, dim3(1), dim3(64), 0, stream, state, n_sweeps, n_samples,
                     block_size);
  return hipGetLastError();
}

template <typename T, int PROP, bool EXT, int MINW>
static hipError_t launch_sj_m(const SJArgs &a, int n_tiles, size_t lds_bytes, hipStream_t stream) {
  if (lds_bytes > 65536) {
    // more than the default 64 KiB of dynamic LDS needs an opt-in per kernel
    const hipError_t e = hipFuncSetAttribute((const void *)sj_sweep_kernel<T, PROP, EXT, MINW>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((sj_sweep_kernel<T, PROP, EXT, MINW>), dim3(n_tiles), dim3(64), lds_bytes, stream, a);
  return hipGetLastError();
}

template <typename T, int PROP, bool EXT>
static hipError_t launch_sj_e(const SJArgs &a, int n_tiles, size_t lds_bytes, hipStream_t stream) {
  if constexpr (PROP == kSJPropLearned)
    if (a.one_wave) return launch_sj_m<T, PROP, EXT, 1>(a, n_tiles, lds_bytes, stream);
  return launch_sj_m<T, PROP, EXT, 2>(a, n_tiles, lds_bytes, stream);
}

template <typename T, int PROP>
static hipError_t launch_sj_t(const SJArgs &a, int n_tiles, size_t lds_bytes, hipStream_t stream) {
  return a.ext ? launch_sj_e<T, PROP, true>(a, n_tiles, lds_bytes, stream)
               : launch_sj_e<T, PROP, false>(a, n_tiles, lds_bytes, stream);
}

template <typename T>
static hipError_t launch_sj_p(const SJArgs &a, int workgroups, size_t lds_bytes, hipStream_t stream) {
  if (lds_bytes > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void *)sj_persistent_kernel<T>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(sj_persistent_kernel<T>, dim3(workgroups), dim3(64), lds_bytes, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sj_ticket_reset_kernel, dim3(1), dim3(64), 0, stream, a.ticket);
  return hipGetLastError();
}

hipError_t launch_sj_persistent(const SJArgs &a, int workgroups, int fp32, size_t lds_bytes, hipStream_t stream) {
  if (a.n_tiles <= 0 || workgroups <= 0) return hipSuccess;
  return fp32 ? launch_sj_p<float>(a, workgroups, lds_bytes, stream)
              : launch_sj_p<double>(a, workgroups, lds_bytes, stream);
}

int sj_persistent_occupancy(int fp32, size_t lds_bytes) {
  int n = 0;
  const hipError_t e =
      fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sj_persistent_kernel<float>, 64, lds_bytes)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sj_persistent_kernel<double>, 64, lds_bytes);
  return e == hipSuccess ? n : 0;
}

namespace {
// One optimiser step of a Slater-Jastrow system on the device (the loop of
// VmcRunner::run_optimization, src/vmc/src/vmc.rs:43-179): the reduced
// sums of the system's NP parameters (columns sj_columns(NP)),
// Optimizer::compute_parameter_update,
// then the parameters the sweep kernel reads (sys->p) and, with the
// exponent scale, the exponents s zeta0_m.
template <int NP>
__global__ void sj_vmc_step_kernel(const double *red, double n_samples, double *params, mole_optimizer *opt,
                                   double *log, SJSysDev *sys, const double *zeta0, double *zeta, int n_basis) {
  __shared__ double scale;
  __shared__ int updated;
  if (threadIdx.x == 0) {
    double so[NP], soe[NP], soo[NP * NP], dp[NP], pars[NP];
    for (int k = 0; k < NP; ++k) {
      so[k] = red[kSJSumO + k];
      soe[k] = red[sj_sum_oe(NP) + k];
      pars[k] = params[k];
      for (int l = k; l < NP; ++l) {
        const int t = k * NP - k * (k - 1) / 2 + (l - k);  // upper-triangle position of (k, l)
        soo[k * NP + l] = red[sj_sum_oo(NP) + t];
        soo[l * NP + k] = red[sj_sum_oo(NP) + t];
      }
    }
    const int rc = optimizer_update<NP>(opt, n_samples, red[kSJSumE], so, soe, soo, pars, dp, nullptr);
    log[0] = red[kSJSumE];
    log[1] = red[kSJSumE2];
    log[2] = red[kSJSumBm];
    log[3] = red[kSJSumBm2];
    log[4] = red[kSJAccepted];
    log[5] = (double)rc;
    log[6] = n_samples;
    log[7] = 0.0;
    updated = rc == 0;
    if (rc == 0)
      for (int k = 0; k < NP; ++k) {
        params[k] = pars[k] + dp[k];
        sys->p[k] = params[k];
      }
    scale = NP == kSJNParam ? params[kSJNParam - 1] : 1.0;
  }
  __syncthreads();
  if (NP == kSJNParam && updated)
    for (int m = threadIdx.x; m < n_basis; m += blockDim.x) zeta[m] = scale * zeta0[m];
}
}  // namespace

hipError_t launch_sj_vmc_step(const double *red, double n_samples, int n_params, double *params, mole_optimizer *opt,
                              double *log, SJSysDev *sys, const double *zeta0, double *zeta, int n_basis,
                              hipStream_t stream) {
  if (n_params == kSJNParam)
    hipLaunchKernelGGL(sj_vmc_step_kernel<kSJNParam>, dim3(1), dim3(64), 0, stream, red, n_samples, params, opt,
                       log, sys, zeta0, zeta, n_basis);
  else
    hipLaunchKernelGGL(sj_vmc_step_kernel<kSJNParam - 1>, dim3(1), dim3(64), 0, stream, red, n_samples, params,
                       opt, log, sys, zeta0, zeta, n_basis);
  return hipGetLastError();
}

hipError_t launch_sj(const SJArgs &a, int n_tiles, int fp32, int proposal, size_t lds_bytes,
                     hipStream_t stream) {
  if (n_tiles <= 0) return hipSuccess;
  if (proposal == kSJPropDiffuse)
    return fp32 ? launch_sj_t<float, kSJPropDiffuse>(a, n_tiles, lds_bytes, stream)
                : launch_sj_t<double, kSJPropDiffuse>(a, n_tiles, lds_bytes, stream);
  if (proposal == kSJPropLearned)
    return fp32 ? launch_sj_t<float, kSJPropLearned>(a, n_tiles, lds_bytes, stream)
                : launch_sj_t<double, kSJPropLearned>(a, n_tiles, lds_bytes, stream);
  return fp32 ? launch_sj_t<float, kSJPropBox>(a, n_tiles, lds_bytes, stream)
              : launch_sj_t<double, kSJPropBox>(a, n_tiles, lds_bytes, stream);
}

}  // namespace mole
