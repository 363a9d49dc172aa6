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

template <int KIND, bool SR>
__global__ void __launch_bounds__(256) sweep2_kernel(Sweep2Args a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.W) return;
  const Sys2 &sp = a.sp;
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

  double acc[kAccColumns2];
#pragma unroll
  for (int k = 0; k < kAccColumns2; ++k) acc[k] = a.acc[(size_t)k * W + w];

  Stream rng;
  rng.init(a.seed, a.sys, (uint32_t)w, a.epoch);
  uint64_t p = a.draw0;
  uint32_t fill = a.block_fill;
  uint32_t sample_idx = a.trace_sample0;
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
      const bool acc_move = ratio > u;
      if (acc_move) {
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
        acc[kAccAccepted] += 1.0;
        mask |= (uint64_t)1 << e;
      }
    }
    if (a.trace_mask && a.trace_sweep0 + s < a.cap_sweeps)
      a.trace_mask[(size_t)w * a.cap_sweeps + a.trace_sweep0 + s] = mask;

    if (s >= a.sample_start) {
      const Local2 L = local2<KIND>(sp, x);
      const double E = L.e_local;
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
      if (SR || a.samp_g) {
        if (a.o_conv == MOLE_O_REFERENCE) {
          // G = psi ./ (psi * dpsi), dpsi = psi dlog; O = G / psi_s with
          // psi_s = (psi*psi)/psi the "Wavefunction value" sample
          const double psi_s = (psi * psi) / psi;
#pragma unroll
          for (int k = 0; k < kNParam2; ++k)
            O[k] = (psi / (psi * (psi * L.dlog[k]))) / psi_s;
        } else {
#pragma unroll
          for (int k = 0; k < kNParam2; ++k) O[k] = L.dlog[k];
        }
      }
      if (SR) {
        int t = 0;
#pragma unroll
        for (int k = 0; k < kNParam2; ++k) {
          acc[kAccSumO + k] += O[k];
          acc[kAccSumOE + k] = fma(O[k], E, acc[kAccSumOE + k]);
#pragma unroll
          for (int l = k; l < kNParam2; ++l, ++t)
            acc[kAccSumOO + t] = fma(O[k], O[l], acc[kAccSumOO + t]);
        }
      }
      if (sample_idx < a.cap_samples) {
        const size_t si = (size_t)w * a.cap_samples + sample_idx;
        if (a.trace_e) a.trace_e[si] = E;
        if (a.samp_psi) a.samp_psi[si] = (psi * psi) / psi;
        if (a.samp_g) {
          const double psi_s = (psi * psi) / psi;
#pragma unroll
          for (int k = 0; k < kNParam2; ++k) a.samp_g[si * kNParam2 + k] = O[k] * psi_s;
        }
      }
      ++sample_idx;
    }
  }

#pragma unroll
  for (int k = 0; k < 6; ++k) a.pos[(size_t)k * W + w] = x[k];
#pragma unroll
  for (int k = 0; k < kAccColumns2; ++k) a.acc[(size_t)k * W + w] = acc[k];
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
__global__ void __launch_bounds__(256) reduce_columns_kernel(const double *acc, int W,
                                                             double *out) {
  __shared__ double sh[256];
  const int col = blockIdx.x;
  const double *p = acc + (size_t)col * W;
  double s = 0.0;
  for (int i = threadIdx.x; i < W; i += 256) s += p[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = sh[0];
}

__global__ void fill_kernel(double *p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_sweep2(const Sweep2Args &a, int block_threads, hipStream_t stream) {
  const dim3 grid((a.W + block_threads - 1) / block_threads);
  const bool sr = true;
  if (a.sp.kind == MOLE_WF_HELIUM_PADE) {
    if (sr) hipLaunchKernelGGL((sweep2_kernel<MOLE_WF_HELIUM_PADE, true>), grid, dim3(block_threads), 0, stream, a);
  } else {
    if (sr) hipLaunchKernelGGL((sweep2_kernel<MOLE_WF_LCAO2_PADE, true>), grid, dim3(block_threads), 0, stream, a);
  }
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

hipError_t launch_reduce_columns(const double *acc, int W, int K, double *out, hipStream_t stream) {
  hipLaunchKernelGGL(reduce_columns_kernel, dim3(K), dim3(256), 0, stream, acc, W, out);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

}  // namespace mole
