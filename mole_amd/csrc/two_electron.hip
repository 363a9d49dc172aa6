// Two-electron walker kernels (He singlet with Pade e-e factor; LCAO
// sigma_g^2 with Pade e-e factor).
//
// One lane owns one walker for the whole launch: positions, the cached
// psi and the distances it was built from, and every accumulator live in
// VGPRs across all sweeps of the launch; HBM is touched once at entry and
// once at exit (positions as [6][W] and accumulators as [column][W], both
// coalesced).  Each move does what MetropolisBox::move_state does
// (src/metropolis/src/metrop.rs:60-96): three Range draws, one Standard
// draw, psi at the proposal, min(psi'^2/psi^2, 1) > u.  psi at the
// current configuration is the cached value of the same expression, so
// every accept decision equals the reference's two-evaluation form bit
// for bit.  After each sampled sweep the local energy
// -1/2 lap(psi)/psi + V (src/operator/src/operator.rs:177-183) and the
// parameter log-derivatives are formed in closed form and folded into the
// per-walker sums that the energy gradient and the SR matrix need
// (src/optimize/src/util.rs:6-46, src/optimize/src/optimizers.rs:191-233).
#include "../../include/mole_gpu.h"
#include "device_common.hpp"
#include "internal.hpp"
#include "sr_math.hpp"

namespace mole {

// Orbital factor of one electron for the LCAO kind:
// phi = sum_I exp(-zeta r_I), spelled as the value expression requires.
__device__ __forceinline__ double lcao_phi(const Sys2 &sp, double x, double y,
                                           double z) {
  double phi = 0.0;
  for (int I = 0; I < sp.n_nuc; ++I) {
    const double r = norm3(x - sp.nuc[I][0], y - sp.nuc[I][1], z - sp.nuc[I][2]);
    const double e = exp_m(-sp.zeta * r);
    phi = (I == 0) ? e : phi + e;
  }
  return phi;
}

__device__ __forceinline__ double pade(double b, double c, double r) {
  return (b * r) / (1.0 + c * r);
}

// Closed-form local quantities of one configuration: E_L, the laplacian
// ratio sum_i lap_i(psi)/psi, grad ln psi and d ln psi/dp.
struct Local2 {
  double e_local;
  double lap_ratio;
  double g[6];
  double dlog[kNParam2];
};

template <int KIND>
__device__ __forceinline__ Local2 local2(const Sys2 &sp, const double *x) {
  Local2 L;
  const double zeta = sp.zeta, b = sp.b, c = sp.c;
  double V = sp.v_nn;
  double gi[2][3], li[2], dz = 0.0;
  for (int e = 0; e < 2; ++e) {
    const double *xe = x + 3 * e;
    double phi = 0.0, gs[3] = {0.0, 0.0, 0.0}, ls = 0.0, zs = 0.0;
    const int nn = (KIND == MOLE_WF_HELIUM_PADE) ? 1 : sp.n_nuc;
    for (int I = 0; I < sp.n_nuc; ++I) {
      const double d0 = xe[0] - sp.nuc[I][0], d1 = xe[1] - sp.nuc[I][1],
                   d2 = xe[2] - sp.nuc[I][2];
      const double r = norm3(d0, d1, d2);
      const double inv = 1.0 / r;
      V = V - sp.charge[I] * inv;
      if (I < nn) {
        // helium kind: phi = e^{-zeta r}, so the ratios need no exponential
        const double ee = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : exp_m(-zeta * r);
        phi += ee;
        const double f = -zeta * ee * inv;
        gs[0] = fma(f, d0, gs[0]);
        gs[1] = fma(f, d1, gs[1]);
        gs[2] = fma(f, d2, gs[2]);
        ls = fma(ee, zeta * zeta - 2.0 * zeta * inv, ls);
        zs = fma(-r, ee, zs);
      }
    }
    const double ip = 1.0 / phi;
    gi[e][0] = gs[0] * ip;
    gi[e][1] = gs[1] * ip;
    gi[e][2] = gs[2] * ip;
    li[e] = ls * ip;
    dz += zs * ip;
  }
  const double d0 = x[0] - x[3], d1 = x[1] - x[4], d2 = x[2] - x[5];
  const double r12 = norm3(d0, d1, d2);
  const double i12 = 1.0 / r12;
  const double den = 1.0 + c * r12;
  const double iden = 1.0 / den;
  const double up = b * iden * iden;
  const double upp = -2.0 * b * c * iden * iden * iden;
  const double J = upp + 2.0 * up * i12 + up * up;
  const double h0 = d0 * i12, h1 = d1 * i12, h2 = d2 * i12;
  const double gd1 = gi[0][0] * h0 + gi[0][1] * h1 + gi[0][2] * h2;
  const double gd2 = gi[1][0] * h0 + gi[1][1] * h1 + gi[1][2] * h2;
  const double k1 = li[0] + J + 2.0 * up * gd1;
  const double k2 = li[1] + J - 2.0 * up * gd2;
  if (sp.hamiltonian == MOLE_H_ELECTRONIC) V = V + i12;
  L.lap_ratio = k1 + k2;
  L.e_local = -0.5 * (k1 + k2) + V;
  L.g[0] = gi[0][0] + up * h0;
  L.g[1] = gi[0][1] + up * h1;
  L.g[2] = gi[0][2] + up * h2;
  L.g[3] = gi[1][0] - up * h0;
  L.g[4] = gi[1][1] - up * h1;
  L.g[5] = gi[1][2] - up * h2;
  L.dlog[0] = dz;
  L.dlog[1] = r12 * iden;
  L.dlog[2] = -b * r12 * r12 * iden * iden;
  return L;
}

template <int KIND>
__device__ __forceinline__ double value2(const Sys2 &sp, const double *x) {
  const double r12 = norm3(x[0] - x[3], x[1] - x[4], x[2] - x[5]);
  if (KIND == MOLE_WF_HELIUM_PADE) {
    const double r1 = norm3(x[0] - sp.nuc[0][0], x[1] - sp.nuc[0][1], x[2] - sp.nuc[0][2]);
    const double r2 = norm3(x[3] - sp.nuc[0][0], x[4] - sp.nuc[0][1], x[5] - sp.nuc[0][2]);
    return exp_m(-sp.zeta * (r1 + r2) + pade(sp.b, sp.c, r12));
  }
  const double phi1 = lcao_phi(sp, x[0], x[1], x[2]);
  const double phi2 = lcao_phi(sp, x[3], x[4], x[5]);
  return (phi1 * phi2) * exp_m(pade(sp.b, sp.c, r12));
}

// Per-walker sample accumulators: energy sums, block means, SR sums, and
// the optional trace / compat-sample writes, in sample order.
struct SampleAcc {
  double acc[kAccColumns2];
  uint32_t fill, sample_idx;

  __device__ __forceinline__ void load(const Sweep2Args &a, int w) {
#pragma unroll
    for (int k = 0; k < kAccColumns2; ++k)
      if (k != kAccAccepted) acc[k] = a.fresh ? 0.0 : a.acc[(size_t)k * a.W + w];
    fill = a.block_fill;
    sample_idx = a.trace_sample0;
  }
  __device__ __forceinline__ void store(const Sweep2Args &a, int w) const {
#pragma unroll
    for (int k = 0; k < kAccColumns2; ++k)
      if (k != kAccAccepted) a.acc[(size_t)k * a.W + w] = acc[k];
  }
  // write: this lane records the sample's trace and compat entries (the
  // sums are kept either way)
  __device__ __forceinline__ void add(const Sweep2Args &a, int w, double E,
                                      const double *dlog, double psi, double inv_block, bool write = true) {
    acc[kAccSumE] += E;
    acc[kAccSumE2] = fma(E, E, acc[kAccSumE2]);
    acc[kAccBlockSum] += E;
    if (++fill == a.block_size) {
      const double bm = acc[kAccBlockSum] * inv_block;
      acc[kAccSumBm] += bm;
      acc[kAccSumBm2] = fma(bm, bm, acc[kAccSumBm2]);
      acc[kAccBlockSum] = 0.0;
      fill = 0;
    }
    double O[kNParam2];
    if (a.o_conv == MOLE_O_REFERENCE) {
      // G = psi ./ (psi * dpsi), dpsi = psi dlog; O = G / psi_s with
      // psi_s = (psi*psi)/psi the "Wavefunction value" sample
      const double psi_s = (psi * psi) / psi;
#pragma unroll
      for (int k = 0; k < kNParam2; ++k) O[k] = (psi / (psi * (psi * dlog[k]))) / psi_s;
    } else {
#pragma unroll
      for (int k = 0; k < kNParam2; ++k) O[k] = dlog[k];
    }
    int t = 0;
#pragma unroll
    for (int k = 0; k < kNParam2; ++k) {
      acc[kAccSumO + k] += O[k];
      acc[kAccSumOE + k] = fma(O[k], E, acc[kAccSumOE + k]);
#pragma unroll
      for (int l = k; l < kNParam2; ++l, ++t)
        acc[kAccSumOO + t] = fma(O[k], O[l], acc[kAccSumOO + t]);
    }
    if (write && sample_idx < a.cap_samples) {
      const size_t si = (size_t)w * a.cap_samples + sample_idx;
      const double psi_s = (psi * psi) / psi;
      if (a.trace_e) a.trace_e[si] = E;
      if (a.samp_psi) a.samp_psi[si] = psi_s;
      if (a.samp_g)
#pragma unroll
        for (int k = 0; k < kNParam2; ++k) a.samp_g[si * kNParam2 + k] = O[k] * psi_s;
    }
    ++sample_idx;
  }
};

// One lane per walker, every stage in the same wave: the general path
// (any number of nuclei up to kMaxNuc2).
template <int KIND>
__global__ void __launch_bounds__(256) sweep2_lane_kernel(Sweep2Args a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.W) return;
  Sys2 sp = a.sp;
  if (a.dparams) {  // parameters updated on the device between launches
    sp.zeta = a.dparams[0];
    sp.b = a.dparams[1];
    sp.c = a.dparams[2];
  }
  const int W = a.W;
  double x[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = a.pos[(size_t)k * W + w];

  // cached pieces of psi at the current configuration
  double r1 = 0.0, r2 = 0.0, phi1 = 0.0, phi2 = 0.0;
  double r12 = norm3(x[0] - x[3], x[1] - x[4], x[2] - x[5]);
  double psi;
  if (KIND == MOLE_WF_HELIUM_PADE) {
    r1 = norm3(x[0] - sp.nuc[0][0], x[1] - sp.nuc[0][1], x[2] - sp.nuc[0][2]);
    r2 = norm3(x[3] - sp.nuc[0][0], x[4] - sp.nuc[0][1], x[5] - sp.nuc[0][2]);
    psi = exp_m(-sp.zeta * (r1 + r2) + pade(sp.b, sp.c, r12));
  } else {
    phi1 = lcao_phi(sp, x[0], x[1], x[2]);
    phi2 = lcao_phi(sp, x[3], x[4], x[5]);
    psi = (phi1 * phi2) * exp_m(pade(sp.b, sp.c, r12));
  }

  SampleAcc sa;
  sa.load(a, w);
  double accepted = a.fresh ? 0.0 : a.acc[(size_t)kAccAccepted * W + w];
  Stream rng;
  rng.init(a.seed, a.sys, (uint32_t)w, a.epoch);
  uint64_t p = a.draw0;
  const double inv_block = 1.0 / (double)a.block_size;

  for (uint32_t s = 0; s < a.n_sweeps; ++s) {
    uint64_t mask = 0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double dx = u64_range(rng.draw(p), a.scale, a.offset);
      const double dy = u64_range(rng.draw(p + 1), a.scale, a.offset);
      const double dz = u64_range(rng.draw(p + 2), a.scale, a.offset);
      const double u = u64_standard(rng.draw(p + 3));
      p += 4;
      const double nx = x[3 * e] + dx, ny = x[3 * e + 1] + dy, nz = x[3 * e + 2] + dz;
      const double n12 = (e == 0) ? norm3(nx - x[3], ny - x[4], nz - x[5])
                                  : norm3(x[0] - nx, x[1] - ny, x[2] - nz);
      double psin, nr = 0.0, nphi = 0.0;
      if (KIND == MOLE_WF_HELIUM_PADE) {
        nr = norm3(nx - sp.nuc[0][0], ny - sp.nuc[0][1], nz - sp.nuc[0][2]);
        const double s12 = (e == 0) ? (nr + r2) : (r1 + nr);
        psin = exp_m(-sp.zeta * s12 + pade(sp.b, sp.c, n12));
      } else {
        nphi = lcao_phi(sp, nx, ny, nz);
        const double jas = exp_m(pade(sp.b, sp.c, n12));
        psin = ((e == 0) ? (nphi * phi2) : (phi1 * nphi)) * jas;
      }
      const double ratio = min_one((psin * psin) / (psi * psi));
      if (ratio > u) {
        x[3 * e] = nx;
        x[3 * e + 1] = ny;
        x[3 * e + 2] = nz;
        r12 = n12;
        psi = psin;
        if (KIND == MOLE_WF_HELIUM_PADE) {
          if (e == 0) r1 = nr; else r2 = nr;
        } else {
          if (e == 0) phi1 = nphi; else phi2 = nphi;
        }
        accepted += 1.0;
        mask |= (uint64_t)1 << e;
      }
    }
    if (a.trace_mask && a.trace_sweep0 + s < a.cap_sweeps)
      a.trace_mask[(size_t)w * a.cap_sweeps + a.trace_sweep0 + s] = mask;
    if (s >= a.sample_start) {
      const Local2 L = local2<KIND>(sp, x);
      sa.add(a, w, L.e_local, L.dlog, psi, inv_block);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) a.pos[(size_t)k * W + w] = x[k];
  sa.store(a, w);
  a.acc[(size_t)kAccAccepted * W + w] = accepted;
}

// Local energy and d ln psi/dp from a configuration whose distances to the
// nuclei (and, for the LCAO kind, the orbital exponentials) are known.
template <int KIND, int NNUC>
__device__ __forceinline__ void local2_cached(const Sys2 &sp, const double *x,
                                              const double (*r)[NNUC],
                                              const double (*ee)[NNUC], double r12,
                                              double &E, double *dlog) {
  const double zeta = sp.zeta, b = sp.b, c = sp.c;
  double V = sp.v_nn;
  double gi[2][3], li[2], dz = 0.0;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    double phi = 0.0, gs0 = 0.0, gs1 = 0.0, gs2 = 0.0, ls = 0.0, zs = 0.0;
#pragma unroll
    for (int I = 0; I < NNUC; ++I) {
      const double d0 = x[3 * e] - sp.nuc[I][0], d1 = x[3 * e + 1] - sp.nuc[I][1],
                   d2 = x[3 * e + 2] - sp.nuc[I][2];
      const double inv = 1.0 / r[e][I];
      V = V - sp.charge[I] * inv;
      const double eI = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : ee[e][I];
      phi += eI;
      const double f = -zeta * eI * inv;
      gs0 = fma(f, d0, gs0);
      gs1 = fma(f, d1, gs1);
      gs2 = fma(f, d2, gs2);
      ls = fma(eI, zeta * zeta - 2.0 * zeta * inv, ls);
      zs = fma(-r[e][I], eI, zs);
    }
    const double ip = 1.0 / phi;
    gi[e][0] = gs0 * ip;
    gi[e][1] = gs1 * ip;
    gi[e][2] = gs2 * ip;
    li[e] = ls * ip;
    dz += zs * ip;
  }
  const double d0 = x[0] - x[3], d1 = x[1] - x[4], d2 = x[2] - x[5];
  const double i12 = 1.0 / r12;
  const double iden = 1.0 / (1.0 + c * r12);
  const double up = b * iden * iden;
  const double upp = -2.0 * b * c * iden * iden * iden;
  const double J = upp + 2.0 * up * i12 + up * up;
  const double h0 = d0 * i12, h1 = d1 * i12, h2 = d2 * i12;
  const double gd1 = gi[0][0] * h0 + gi[0][1] * h1 + gi[0][2] * h2;
  const double gd2 = gi[1][0] * h0 + gi[1][1] * h1 + gi[1][2] * h2;
  const double k1 = li[0] + J + 2.0 * up * gd1;
  const double k2 = li[1] + J - 2.0 * up * gd2;
  if (sp.hamiltonian == MOLE_H_ELECTRONIC) V = V + i12;
  E = -0.5 * (k1 + k2) + V;
  dlog[0] = dz;
  dlog[1] = r12 * iden;
  dlog[2] = -b * r12 * r12 * iden * iden;
}

// Wave-specialised sweep: a workgroup of four waves owns 64 walkers.
//   wave 0      the Metropolis chain (proposal, psi', accept, state update);
//   waves 1, 2  Philox draws of chunk c+1 (electron 0 / electron 1) into LDS;
//   wave 3      local energy, d ln psi/dp and accumulators of chunk c-1,
//               from state snapshots the chain leaves in LDS.
// Chunks are T sweeps; every buffer is double-buffered and the four roles
// meet at one barrier per chunk, so the chain -- the only serial part --
// runs alone on its SIMD and its arithmetic is exactly that of the
// one-wave kernel.
template <int KIND, int NNUC, int T>
__global__ void __launch_bounds__(256) sweep2_ws_kernel(Sweep2Args a) {
  constexpr int SNAP = 6 + 4 * NNUC + 2;  // x, r[e][I], e[e][I], r12, psi
  // one block with the draws first: their LDS addresses then start at 0, so
  // the eight draws of a sweep are one base register plus immediate offsets
  struct Lds {
    double rng[2][T][2][4][64];
    double snap[2][T][SNAP][64];
  };
  __shared__ Lds lds;
  auto &rng_lds = lds.rng;
  auto &snap_lds = lds.snap;
  const int lane = threadIdx.x & 63;
  const int role = threadIdx.x >> 6;
  const int w = blockIdx.x * 64 + lane;
  const bool live = w < a.W;
  const int W = a.W;
  Sys2 sp = a.sp;
  if (a.dparams) {  // parameters updated on the device between launches
    sp.zeta = a.dparams[0];
    sp.b = a.dparams[1];
    sp.c = a.dparams[2];
  }
  const uint32_t nchunks = (a.n_sweeps + T - 1) / T;
  const double inv_block = 1.0 / (double)a.block_size;

  // chain state
  double x[6];
  double r[2][NNUC], ee[2][NNUC], phi[2];
  double r12 = 0.0, psi = 1.0, accepted = 0.0;
  // producer state
  Stream rng;
  // sampler state
  SampleAcc sa;

  if (role == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = live ? a.pos[(size_t)k * W + w] : 0.0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      phi[e] = 0.0;
#pragma unroll
      for (int I = 0; I < NNUC; ++I) {
        r[e][I] = norm3(x[3 * e] - sp.nuc[I][0], x[3 * e + 1] - sp.nuc[I][1],
                        x[3 * e + 2] - sp.nuc[I][2]);
        ee[e][I] = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : exp_m(-sp.zeta * r[e][I]);
        phi[e] = (I == 0) ? ee[e][I] : phi[e] + ee[e][I];
      }
    }
    r12 = norm3(x[0] - x[3], x[1] - x[4], x[2] - x[5]);
    if (KIND == MOLE_WF_HELIUM_PADE) {
      psi = exp_m(-sp.zeta * (r[0][0] + r[1][0]) + pade(sp.b, sp.c, r12));
    } else {
      psi = (phi[0] * phi[1]) * exp_m(pade(sp.b, sp.c, r12));
    }
    accepted = (live && !a.fresh) ? a.acc[(size_t)kAccAccepted * W + w] : 0.0;
  } else if (role == 3) {
    if (live) sa.load(a, w);
  } else {
    rng.init(a.seed, a.sys, (uint32_t)w, a.epoch);
  }

  // producers: draws of chunk cc for electron e = role - 1
  auto produce = [&](uint32_t cc) {
    const int e = role - 1;
    const int buf = cc & 1;
    const uint32_t s0 = cc * T;
    for (int t = 0; t < T; ++t) {
      if (s0 + t >= a.n_sweeps) break;
      const uint64_t p = a.draw0 + ((uint64_t)(s0 + t) * 2 + e) * 4;
      rng_lds[buf][t][e][0][lane] = u64_range(rng.draw(p), a.scale, a.offset);
      rng_lds[buf][t][e][1][lane] = u64_range(rng.draw(p + 1), a.scale, a.offset);
      rng_lds[buf][t][e][2][lane] = u64_range(rng.draw(p + 2), a.scale, a.offset);
      rng_lds[buf][t][e][3][lane] = u64_standard(rng.draw(p + 3));
    }
  };

  if (role == 1 || role == 2) {
    if (nchunks > 0) produce(0);
  }
  __syncthreads();

  for (uint32_t c = 0; c <= nchunks; ++c) {
    const int buf = c & 1;
    const uint32_t s0 = c * T;
    if (role == 0 && c < nchunks) {
      for (int t = 0; t < T; ++t) {
        const uint32_t s = s0 + t;
        if (s >= a.n_sweeps) break;
        // the sweep's eight draws read from LDS together, ahead of the arithmetic
        double dr[2][4];
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int k = 0; k < 4; ++k) dr[e][k] = rng_lds[buf][t][e][k][lane];
        uint64_t mask = 0;
        // electron-local parts of both moves first (trial position, distances to
        // the nuclei, orbital values: none depends on the other electron's
        // decision), then each move's coupled part (r12, Jastrow, ratio) in
        // order; the updates are selects.  Every value is computed by the same
        // operations from the same operands as move-by-move code, so the chain
        // is bit-identical while its two halves overlap in the pipeline.
        double nx[2][3], nr[2][NNUC], ne[2][NNUC], nphi[2], uu[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          nx[e][0] = x[3 * e] + dr[e][0];
          nx[e][1] = x[3 * e + 1] + dr[e][1];
          nx[e][2] = x[3 * e + 2] + dr[e][2];
          uu[e] = dr[e][3];
          nphi[e] = 0.0;
#pragma unroll
          for (int I = 0; I < NNUC; ++I) {
            nr[e][I] = norm3(nx[e][0] - sp.nuc[I][0], nx[e][1] - sp.nuc[I][1], nx[e][2] - sp.nuc[I][2]);
            ne[e][I] = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : exp_m(-sp.zeta * nr[e][I]);
            nphi[e] = (I == 0) ? ne[e][I] : nphi[e] + ne[e][I];
          }
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const double n12 = (e == 0) ? norm3(nx[0][0] - x[3], nx[0][1] - x[4], nx[0][2] - x[5])
                                      : norm3(x[0] - nx[1][0], x[1] - nx[1][1], x[2] - nx[1][2]);
          double psin;
          if (KIND == MOLE_WF_HELIUM_PADE) {
            const double s12 = (e == 0) ? (nr[0][0] + r[1][0]) : (r[0][0] + nr[1][0]);
            psin = exp_m(-sp.zeta * s12 + pade(sp.b, sp.c, n12));
          } else {
            const double jas = exp_m(pade(sp.b, sp.c, n12));
            psin = ((e == 0) ? (nphi[0] * phi[1]) : (phi[0] * nphi[1])) * jas;
          }
          const double ratio = min_one((psin * psin) / (psi * psi));
          const bool acc = ratio > uu[e];
          x[3 * e] = sel(acc, nx[e][0], x[3 * e]);
          x[3 * e + 1] = sel(acc, nx[e][1], x[3 * e + 1]);
          x[3 * e + 2] = sel(acc, nx[e][2], x[3 * e + 2]);
#pragma unroll
          for (int I = 0; I < NNUC; ++I) {
            r[e][I] = sel(acc, nr[e][I], r[e][I]);
            ee[e][I] = sel(acc, ne[e][I], ee[e][I]);
          }
          phi[e] = sel(acc, nphi[e], phi[e]);
          r12 = sel(acc, n12, r12);
          psi = sel(acc, psin, psi);
          accepted += acc ? 1.0 : 0.0;
          mask |= acc ? ((uint64_t)1 << e) : 0;
        }
        if (live && a.trace_mask && a.trace_sweep0 + s < a.cap_sweeps)
          a.trace_mask[(size_t)w * a.cap_sweeps + a.trace_sweep0 + s] = mask;
        double *sn = &snap_lds[buf][t][0][0];
#pragma unroll
        for (int k = 0; k < 6; ++k) sn[k * 64 + lane] = x[k];
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int I = 0; I < NNUC; ++I) {
            sn[(6 + e * NNUC + I) * 64 + lane] = r[e][I];
            sn[(6 + 2 * NNUC + e * NNUC + I) * 64 + lane] = ee[e][I];
          }
        sn[(6 + 4 * NNUC) * 64 + lane] = r12;
        sn[(7 + 4 * NNUC) * 64 + lane] = psi;
      }
    } else if ((role == 1 || role == 2) && c + 1 < nchunks) {
      produce(c + 1);
    } else if (role == 3 && c >= 1) {
      const int pb = (c - 1) & 1;
      const uint32_t p0 = (c - 1) * T;
      for (int t = 0; t < T; ++t) {
        const uint32_t s = p0 + t;
        if (s >= a.n_sweeps) break;
        if (s < a.sample_start) continue;
        const double *sn = &snap_lds[pb][t][0][0];
        double xs[6], rs[2][NNUC], es[2][NNUC];
#pragma unroll
        for (int k = 0; k < 6; ++k) xs[k] = sn[k * 64 + lane];
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int I = 0; I < NNUC; ++I) {
            rs[e][I] = sn[(6 + e * NNUC + I) * 64 + lane];
            es[e][I] = sn[(6 + 2 * NNUC + e * NNUC + I) * 64 + lane];
          }
        const double r12s = sn[(6 + 4 * NNUC) * 64 + lane];
        double E, dlog[kNParam2];
        local2_cached<KIND, NNUC>(sp, xs, rs, es, r12s, E, dlog);
        const double psis = sn[(7 + 4 * NNUC) * 64 + lane];
        if (live) sa.add(a, w, E, dlog, psis, inv_block);
      }
    }
    __syncthreads();
  }

  if (role == 0 && live) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a.pos[(size_t)k * W + w] = x[k];
    a.acc[(size_t)kAccAccepted * W + w] = accepted;
  } else if (role == 3 && live) {
    sa.store(a, w);
  }
}

// ---- four lanes per walker ---------------------------------------------------
//
// A quad of lanes owns one walker; every lane holds the walker's state.  Per
// sweep, lane q
//   - draws u64s p+2q and p+2q+1 (one Philox block when p is even): lanes 0-1
//     hold electron 0's (dx, dy | dz, u), lanes 2-3 electron 1's;
//   - forms electron q/2's trial position and its distance and orbital
//     exponential to nucleus q%2 (the electron-local part of the moves);
//   - forms one coupled part: lane 0 electron 0's (r12', Jastrow / psi'),
//     lane 1 electron 1's if electron 0's move is rejected, lanes 2-3 if it
//     is accepted;
// and the lanes exchange these by DPP quad permutations (register moves, no
// LDS, no barrier).  The remaining chain -- two ratios and the selects of
// the state update -- runs on all four lanes alike.  Every value is the
// same operation on the same operands as in the one-lane chain, so the
// trajectories are bit for bit the other kernels'.  After a sampled sweep
// the local energy is split the same way: lane q divides by r_{q/2, q%2},
// then by phi_0 (lane 0), r12 (lane 1), phi_1 (lane 2), 1 + c r12 (lane 3),
// with local2_cached's operations in its order.  With 4 lanes per walker,
// 16,384 walkers occupy all 1,024 SIMDs of the chip.

constexpr int kQuadXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kQuadLane0 = 0x00, kQuadLane1 = 0x55, kQuadLane2 = 0xAA, kQuadLane3 = 0xFF;

// v of the lane the quad permutation CTRL names (two 32-bit DPP moves)
template <int CTRL>
__device__ __forceinline__ double quad_mov(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// coordinate k of electron 1 (e1) or electron 0
__device__ __forceinline__ double xe_sel(bool e1, const double *x, int k) { return e1 ? x[3 + k] : x[k]; }

template <int KIND, int NNUC>
__global__ void __launch_bounds__(64) sweep2_quad_kernel(Sweep2Args a) {
  static_assert(NNUC == 1 || NNUC == 2, "one lane per (electron, nucleus) pair");
  const int q = threadIdx.x & 3;
  const int w = (int)((blockIdx.x * 64u + threadIdx.x) >> 2);
  const bool live = w < a.W;
  const bool lead = live && q == 0;  // writes the walker's results
  const int W = a.W;
  Sys2 sp = a.sp;
  if (a.dparams) {  // parameters updated on the device between launches
    sp.zeta = a.dparams[0];
    sp.b = a.dparams[1];
    sp.c = a.dparams[2];
  }
  const bool me = (q >> 1) != 0;                 // this lane's electron is 1
  const bool mi = NNUC == 2 && (q & 1) != 0;     // this lane's nucleus is 1
  const double *Rq = mi ? sp.nuc[NNUC - 1] : sp.nuc[0];
  const double inv_block = 1.0 / (double)a.block_size;

  // the walker's state, on every lane of the quad
  double x[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = live ? a.pos[(size_t)k * W + w] : 0.0;
  double r[2][NNUC], ee[2][NNUC], phi[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    phi[e] = 0.0;
#pragma unroll
    for (int I = 0; I < NNUC; ++I) {
      r[e][I] = norm3(x[3 * e] - sp.nuc[I][0], x[3 * e + 1] - sp.nuc[I][1], x[3 * e + 2] - sp.nuc[I][2]);
      ee[e][I] = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : exp_m(-sp.zeta * r[e][I]);
      phi[e] = (I == 0) ? ee[e][I] : phi[e] + ee[e][I];
    }
  }
  double r12 = norm3(x[0] - x[3], x[1] - x[4], x[2] - x[5]);
  double psi = (KIND == MOLE_WF_HELIUM_PADE) ? exp_m(-sp.zeta * (r[0][0] + r[1][0]) + pade(sp.b, sp.c, r12))
                                             : (phi[0] * phi[1]) * exp_m(pade(sp.b, sp.c, r12));
  double accepted = (live && !a.fresh) ? a.acc[(size_t)kAccAccepted * W + w] : 0.0;
  SampleAcc sa;
  if (live) {
    sa.load(a, w);
  } else {
#pragma unroll
    for (int k = 0; k < kAccColumns2; ++k) sa.acc[k] = 0.0;
    sa.fill = a.block_fill;
    sa.sample_idx = a.trace_sample0;
  }
  Stream rng;
  rng.init(a.seed, a.sys, (uint32_t)w, a.epoch);
  uint64_t p = a.draw0;

  for (uint32_t s = 0; s < a.n_sweeps; ++s) {
    // ---- draws: electron q/2's (dx, dy) on its even lane, (dz, u) on its odd one
    const uint64_t da = rng.draw(p + 2 * (uint64_t)q), db = rng.draw(p + 2 * (uint64_t)q + 1);
    p += 8;
    const bool odd = (q & 1) != 0;
    const double va = u64_range(da, a.scale, a.offset);
    const double vb = odd ? u64_standard(db) : u64_range(db, a.scale, a.offset);
    const double pa = quad_mov<kQuadXor1>(va), pb = quad_mov<kQuadXor1>(vb);
    const double dxe = odd ? pa : va, dye = odd ? pb : vb, dze = odd ? va : pa, ue = odd ? vb : pb;
    // ---- electron-local part of this lane's (electron, nucleus)
    const double xe0 = me ? x[3] : x[0], xe1 = me ? x[4] : x[1], xe2 = me ? x[5] : x[2];
    const double nxe0 = xe0 + dxe, nxe1 = xe1 + dye, nxe2 = xe2 + dze;
    const double nrq = norm3(nxe0 - Rq[0], nxe1 - Rq[1], nxe2 - Rq[2]);
    const double neq = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : exp_m(-sp.zeta * nrq);
    // this electron's nuclei (the partner lane holds the other nucleus)
    double nre[NNUC], nee[NNUC];
    if constexpr (NNUC == 2) {
      const double nrp = quad_mov<kQuadXor1>(nrq), nep = quad_mov<kQuadXor1>(neq);
      nre[0] = mi ? nrp : nrq;
      nre[1] = mi ? nrq : nrp;
      nee[0] = mi ? nep : neq;
      nee[1] = mi ? neq : nep;
    } else {
      nre[0] = nrq;
      nee[0] = neq;
    }
    double nphie = nee[0];
    if constexpr (NNUC == 2) nphie = nphie + nee[1];
    // the other electron's, from the lane pair (q ^ 2)
    double nx[2][3], nr[2][NNUC], ne[2][NNUC], nphi[2], uu[2];
    {
      const double o0 = quad_mov<kQuadXor2>(nxe0), o1 = quad_mov<kQuadXor2>(nxe1), o2 = quad_mov<kQuadXor2>(nxe2);
      const double ou = quad_mov<kQuadXor2>(ue), ophi = quad_mov<kQuadXor2>(nphie);
      nx[0][0] = me ? o0 : nxe0;
      nx[0][1] = me ? o1 : nxe1;
      nx[0][2] = me ? o2 : nxe2;
      nx[1][0] = me ? nxe0 : o0;
      nx[1][1] = me ? nxe1 : o1;
      nx[1][2] = me ? nxe2 : o2;
      uu[0] = me ? ou : ue;
      uu[1] = me ? ue : ou;
      nphi[0] = me ? ophi : nphie;
      nphi[1] = me ? nphie : ophi;
#pragma unroll
      for (int I = 0; I < NNUC; ++I) {
        const double orr = quad_mov<kQuadXor2>(nre[I]), oee = quad_mov<kQuadXor2>(nee[I]);
        nr[0][I] = me ? orr : nre[I];
        nr[1][I] = me ? nre[I] : orr;
        ne[0][I] = me ? oee : nee[I];
        ne[1][I] = me ? nee[I] : oee;
      }
    }
    // ---- coupled parts: lane 0 electron 0's move; lane 1 electron 1's after
    // a rejected, lanes 2-3 after an accepted move of electron 0
    const bool c0 = q == 0, c1 = q == 1;
    const double A0 = c1 ? x[0] : nx[0][0], A1 = c1 ? x[1] : nx[0][1], A2 = c1 ? x[2] : nx[0][2];
    const double B0 = c0 ? x[3] : nx[1][0], B1 = c0 ? x[4] : nx[1][1], B2 = c0 ? x[5] : nx[1][2];
    const double n12c = norm3(A0 - B0, A1 - B1, A2 - B2);
    double valc;
    if constexpr (KIND == MOLE_WF_HELIUM_PADE) {
      const double s12 = c0 ? (nr[0][0] + r[1][0]) : (c1 ? (r[0][0] + nr[1][0]) : (nr[0][0] + nr[1][0]));
      valc = exp_m(-sp.zeta * s12 + pade(sp.b, sp.c, n12c));
    } else {
      valc = exp_m(pade(sp.b, sp.c, n12c));
    }
    const double n12_0 = quad_mov<kQuadLane0>(n12c), n12_r = quad_mov<kQuadLane1>(n12c),
                 n12_a = quad_mov<kQuadLane2>(n12c);
    const double v_0 = quad_mov<kQuadLane0>(valc), v_r = quad_mov<kQuadLane1>(valc), v_a = quad_mov<kQuadLane2>(valc);
    // ---- the chain: electron 0, then electron 1 (every lane)
    uint64_t mask = 0;
    {
      const double psin = (KIND == MOLE_WF_HELIUM_PADE) ? v_0 : (nphi[0] * phi[1]) * v_0;
      const double ratio = min_one((psin * psin) / (psi * psi));
      const bool acc = ratio > uu[0];
      x[0] = sel(acc, nx[0][0], x[0]);
      x[1] = sel(acc, nx[0][1], x[1]);
      x[2] = sel(acc, nx[0][2], x[2]);
#pragma unroll
      for (int I = 0; I < NNUC; ++I) {
        r[0][I] = sel(acc, nr[0][I], r[0][I]);
        ee[0][I] = sel(acc, ne[0][I], ee[0][I]);
      }
      phi[0] = sel(acc, nphi[0], phi[0]);
      r12 = sel(acc, n12_0, r12);
      psi = sel(acc, psin, psi);
      accepted += acc ? 1.0 : 0.0;
      mask |= acc ? 1ull : 0ull;
      // electron 1's coupled part for this outcome
      const double n12_1 = acc ? n12_a : n12_r;
      const double v_1 = acc ? v_a : v_r;
      const double psin1 = (KIND == MOLE_WF_HELIUM_PADE) ? v_1 : (phi[0] * nphi[1]) * v_1;
      const double ratio1 = min_one((psin1 * psin1) / (psi * psi));
      const bool acc1 = ratio1 > uu[1];
      x[3] = sel(acc1, nx[1][0], x[3]);
      x[4] = sel(acc1, nx[1][1], x[4]);
      x[5] = sel(acc1, nx[1][2], x[5]);
#pragma unroll
      for (int I = 0; I < NNUC; ++I) {
        r[1][I] = sel(acc1, nr[1][I], r[1][I]);
        ee[1][I] = sel(acc1, ne[1][I], ee[1][I]);
      }
      phi[1] = sel(acc1, nphi[1], phi[1]);
      r12 = sel(acc1, n12_1, r12);
      psi = sel(acc1, psin1, psi);
      accepted += acc1 ? 1.0 : 0.0;
      mask |= acc1 ? 2ull : 0ull;
    }
    if (lead && a.trace_mask && a.trace_sweep0 + s < a.cap_sweeps)
      a.trace_mask[(size_t)w * a.cap_sweeps + a.trace_sweep0 + s] = mask;
    if (s < a.sample_start) continue;
    // ---- local energy and d ln psi/dp (local2_cached's operations, split over the quad)
    const double zeta = sp.zeta, b = sp.b, c = sp.c;
    double inv[2][NNUC];
    {
      const double rq = me ? (mi ? r[1][NNUC - 1] : r[1][0]) : (mi ? r[0][NNUC - 1] : r[0][0]);
      const double iq = 1.0 / rq;
      double ie[NNUC];
      if constexpr (NNUC == 2) {
        const double ip_ = quad_mov<kQuadXor1>(iq);
        ie[0] = mi ? ip_ : iq;
        ie[1] = mi ? iq : ip_;
      } else {
        ie[0] = iq;
      }
#pragma unroll
      for (int I = 0; I < NNUC; ++I) {
        const double oi = quad_mov<kQuadXor2>(ie[I]);
        inv[0][I] = me ? oi : ie[I];
        inv[1][I] = me ? ie[I] : oi;
      }
    }
    // this lane's electron: the sums over nuclei
    double phis = 0.0, gs0 = 0.0, gs1 = 0.0, gs2 = 0.0, ls = 0.0, zs = 0.0;
#pragma unroll
    for (int I = 0; I < NNUC; ++I) {
      const double d0 = xe_sel(me, x, 0) - sp.nuc[I][0], d1 = xe_sel(me, x, 1) - sp.nuc[I][1],
                   d2 = xe_sel(me, x, 2) - sp.nuc[I][2];
      const double iv = me ? inv[1][I] : inv[0][I];
      const double eI = (KIND == MOLE_WF_HELIUM_PADE) ? 1.0 : (me ? ee[1][I] : ee[0][I]);
      phis += eI;
      const double f = -zeta * eI * iv;
      gs0 = fma(f, d0, gs0);
      gs1 = fma(f, d1, gs1);
      gs2 = fma(f, d2, gs2);
      ls = fma(eI, zeta * zeta - 2.0 * zeta * iv, ls);
      zs = fma(-(me ? r[1][I] : r[0][I]), eI, zs);
    }
    // one division per lane: 1/phi_0, 1/r12, 1/phi_1, 1/(1 + c r12)
    const double dv = (q & 1) ? ((q == 1) ? r12 : (1.0 + c * r12)) : phis;
    const double qv = 1.0 / dv;
    const double ip0 = quad_mov<kQuadLane0>(qv), i12 = quad_mov<kQuadLane1>(qv), ip1 = quad_mov<kQuadLane2>(qv),
                 iden = quad_mov<kQuadLane3>(qv);
    const double ogs0 = quad_mov<kQuadXor2>(gs0), ogs1 = quad_mov<kQuadXor2>(gs1), ogs2 = quad_mov<kQuadXor2>(gs2);
    const double ols = quad_mov<kQuadXor2>(ls), ozs = quad_mov<kQuadXor2>(zs);
    double E, dlog[kNParam2];
    {
      double V = sp.v_nn;
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int I = 0; I < NNUC; ++I) V = V - sp.charge[I] * inv[e][I];
      double gi[2][3], li[2], dz = 0.0;
      const double g0[3] = {me ? ogs0 : gs0, me ? ogs1 : gs1, me ? ogs2 : gs2};
      const double g1[3] = {me ? gs0 : ogs0, me ? gs1 : ogs1, me ? gs2 : ogs2};
      const double ls0 = me ? ols : ls, ls1 = me ? ls : ols, zs0 = me ? ozs : zs, zs1 = me ? zs : ozs;
      gi[0][0] = g0[0] * ip0;
      gi[0][1] = g0[1] * ip0;
      gi[0][2] = g0[2] * ip0;
      li[0] = ls0 * ip0;
      dz += zs0 * ip0;
      gi[1][0] = g1[0] * ip1;
      gi[1][1] = g1[1] * ip1;
      gi[1][2] = g1[2] * ip1;
      li[1] = ls1 * ip1;
      dz += zs1 * ip1;
      const double d0 = x[0] - x[3], d1 = x[1] - x[4], d2 = x[2] - x[5];
      const double up = b * iden * iden;
      const double upp = -2.0 * b * c * iden * iden * iden;
      const double J = upp + 2.0 * up * i12 + up * up;
      const double h0 = d0 * i12, h1 = d1 * i12, h2 = d2 * i12;
      const double gd1 = gi[0][0] * h0 + gi[0][1] * h1 + gi[0][2] * h2;
      const double gd2 = gi[1][0] * h0 + gi[1][1] * h1 + gi[1][2] * h2;
      const double k1 = li[0] + J + 2.0 * up * gd1;
      const double k2 = li[1] + J - 2.0 * up * gd2;
      if (sp.hamiltonian == MOLE_H_ELECTRONIC) V = V + i12;
      E = -0.5 * (k1 + k2) + V;
      dlog[0] = dz;
      dlog[1] = r12 * iden;
      dlog[2] = -b * r12 * r12 * iden * iden;
    }
    sa.add(a, w, E, dlog, psi, inv_block, lead);
  }
  if (lead) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a.pos[(size_t)k * W + w] = x[k];
    sa.store(a, w);
    a.acc[(size_t)kAccAccepted * W + w] = accepted;
  }
}

__global__ void init2_kernel(double *pos, int W, uint64_t seed, uint32_t sys,
                             uint32_t epoch) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  Stream rng;
  rng.init(seed, sys, (uint32_t)w, epoch);
  // Range(-1, 1): scale 2, offset -1 - 2
#pragma unroll
  for (int k = 0; k < 6; ++k) pos[(size_t)k * W + w] = u64_range(rng.draw(k), 2.0, -3.0);
}

template <int KIND>
__global__ void eval2_kernel(Eval2Args a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double x[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = a.cfg[i * 6 + k];
  const double psi = value2<KIND>(a.sp, x);
  const Local2 L = local2<KIND>(a.sp, x);
  if (a.value) a.value[i] = psi;
  if (a.grad)
    for (int k = 0; k < 6; ++k) a.grad[i * 6 + k] = psi * L.g[k];
  if (a.lap) a.lap[i] = psi * L.lap_ratio;
  if (a.e_local) a.e_local[i] = L.e_local;
  if (a.pgrad)
    for (int k = 0; k < kNParam2; ++k) a.pgrad[i * kNParam2 + k] = psi * L.dlog[k];
}

// Column sums of a [K][W] array, one workgroup per column, fixed order.
// Column max_col, if in range, is reduced with max instead of +.
__global__ void __launch_bounds__(256) reduce_columns_kernel(const double *acc, int W,
                                                             double *out, int max_first, int max_count) {
  __shared__ double sh[256];
  const int col = blockIdx.x;
  const bool mx = col >= max_first && col < max_first + max_count;
  const double *p = acc + (size_t)col * W;
  double s = 0.0;
  // loads in batches of 16 (all in flight together), added in index order
  int i = threadIdx.x;
  for (; i + 15 * 256 < W; i += 16 * 256) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p[i + q * 256];
#pragma unroll
    for (int q = 0; q < 16; ++q) s = mx ? fmax(s, v[q]) : s + v[q];
  }
  for (; i < W; i += 256) s = mx ? fmax(s, p[i]) : s + p[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      sh[threadIdx.x] = mx ? fmax(sh[threadIdx.x], sh[threadIdx.x + off]) : sh[threadIdx.x] + sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = sh[0];
}

// One workgroup per block: per walker the block's samples summed in order
// times 1/B, then summed over walkers (strided per thread, LDS tree).
__global__ void __launch_bounds__(256) block_records_kernel(const double *trace_e, int W, size_t cap, size_t B,
                                                            size_t first, double *out) {
  __shared__ double s1[256], s2[256];
  const size_t b = first + blockIdx.x;
  const double inv = 1.0 / (double)B;
  double a1 = 0.0, a2 = 0.0;
  for (int w = threadIdx.x; w < W; w += 256) {
    const double *e = trace_e + (size_t)w * cap + b * B;
    double s = 0.0;
    for (size_t k = 0; k < B; ++k) s += e[k];
    const double bm = s * inv;
    a1 += bm;
    a2 = fma(bm, bm, a2);
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s1[threadIdx.x] += s1[threadIdx.x + off];
      s2[threadIdx.x] += s2[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = s1[0];
    out[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ void fill_kernel(double *p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

bool sweep2_wavespec_eligible(const Sys2 &sp) {
  return (sp.kind == MOLE_WF_HELIUM_PADE) ? sp.n_nuc == 1 : sp.n_nuc <= 2;
}

hipError_t launch_sweep2(const Sweep2Args &a, int variant, int block_threads, hipStream_t stream) {
  const bool he = a.sp.kind == MOLE_WF_HELIUM_PADE;
  const int nn = a.sp.n_nuc;
  if (variant == MOLE_VARIANT_QUAD && sweep2_wavespec_eligible(a.sp)) {
    // four lanes per walker, 16 walkers per one-wave workgroup
    const dim3 grid((unsigned)((4 * (size_t)a.W + 63) / 64)), blk(64);
    if (he)
      hipLaunchKernelGGL((sweep2_quad_kernel<MOLE_WF_HELIUM_PADE, 1>), grid, blk, 0, stream, a);
    else if (nn == 1)
      hipLaunchKernelGGL((sweep2_quad_kernel<MOLE_WF_LCAO2_PADE, 1>), grid, blk, 0, stream, a);
    else
      hipLaunchKernelGGL((sweep2_quad_kernel<MOLE_WF_LCAO2_PADE, 2>), grid, blk, 0, stream, a);
    return hipGetLastError();
  }
  if (variant != MOLE_VARIANT_LANE && sweep2_wavespec_eligible(a.sp)) {
    // wave-specialised path: 64 walkers per 256-thread workgroup
    const dim3 grid((a.W + 63) / 64), blk(256);
    constexpr int T = 6;
    if (he)
      hipLaunchKernelGGL((sweep2_ws_kernel<MOLE_WF_HELIUM_PADE, 1, T>), grid, blk, 0, stream, a);
    else if (nn == 1)
      hipLaunchKernelGGL((sweep2_ws_kernel<MOLE_WF_LCAO2_PADE, 1, T>), grid, blk, 0, stream, a);
    else
      hipLaunchKernelGGL((sweep2_ws_kernel<MOLE_WF_LCAO2_PADE, 2, T>), grid, blk, 0, stream, a);
    return hipGetLastError();
  }
  const int bt = block_threads ? block_threads : 64;
  const dim3 grid((a.W + bt - 1) / bt);
  if (he)
    hipLaunchKernelGGL((sweep2_lane_kernel<MOLE_WF_HELIUM_PADE>), grid, dim3(bt), 0, stream, a);
  else
    hipLaunchKernelGGL((sweep2_lane_kernel<MOLE_WF_LCAO2_PADE>), grid, dim3(bt), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_init2(double *pos, int W, uint64_t seed, uint32_t sys, uint32_t epoch,
                        hipStream_t stream) {
  hipLaunchKernelGGL(init2_kernel, dim3((W + 255) / 256), dim3(256), 0, stream, pos, W, seed, sys,
                     epoch);
  return hipGetLastError();
}

hipError_t launch_eval2(const Eval2Args &a, hipStream_t stream) {
  const dim3 grid((unsigned)((a.n + 127) / 128));
  if (a.sp.kind == MOLE_WF_HELIUM_PADE)
    hipLaunchKernelGGL((eval2_kernel<MOLE_WF_HELIUM_PADE>), grid, dim3(128), 0, stream, a);
  else
    hipLaunchKernelGGL((eval2_kernel<MOLE_WF_LCAO2_PADE>), grid, dim3(128), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_reduce_columns(const double *acc, int W, int K, double *out, int max_first, int max_count,
                                 hipStream_t stream) {
  hipLaunchKernelGGL(reduce_columns_kernel, dim3(K), dim3(256), 0, stream, acc, W, out, max_first, max_count);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

namespace {
__global__ void vmc_step_kernel(const double *red, double n_samples, double *params, mole_optimizer *opt,
                                double *log) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  constexpr int np = kNParam2;
  double so[np], soe[np], soo[np * np], dp[np], pars[np];
  int t = 0;
  for (int k = 0; k < np; ++k) {
    so[k] = red[kAccSumO + k];
    soe[k] = red[kAccSumOE + k];
    pars[k] = params[k];
    for (int l = k; l < np; ++l, ++t) {
      soo[k * np + l] = red[kAccSumOO + t];
      soo[l * np + k] = red[kAccSumOO + t];
    }
  }
  // the parameter count is fixed here (mole_optimize_with checks opt->n_params)
  const int rc = optimizer_update<np>(opt, n_samples, red[kAccSumE], so, soe, soo, pars, dp, nullptr);
  log[0] = red[kAccSumE];
  log[1] = red[kAccSumE2];
  log[2] = red[kAccSumBm];
  log[3] = red[kAccSumBm2];
  log[4] = red[kAccAccepted];
  log[5] = (double)rc;
  log[6] = n_samples;
  log[7] = 0.0;
  if (rc == 0)
    for (int k = 0; k < np; ++k) params[k] = pars[k] + dp[k];
}
}  // namespace

hipError_t launch_block_records(const double *trace_e, int W, size_t cap, size_t B, size_t first, size_t k,
                                double *out, hipStream_t stream) {
  hipLaunchKernelGGL(block_records_kernel, dim3((unsigned)k), dim3(256), 0, stream, trace_e, W, cap, B, first, out);
  return hipGetLastError();
}

hipError_t launch_vmc_step(const double *red, double n_samples, double *params, mole_optimizer *opt, double *log,
                           hipStream_t stream) {
  hipLaunchKernelGGL(vmc_step_kernel, dim3(1), dim3(64), 0, stream, red, n_samples, params, opt, log);
  return hipGetLastError();
}

}  // namespace mole
