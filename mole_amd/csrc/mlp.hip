// Walker MLP kernels: Xavier initialisation, bf16 fragment packing, the
// training step (two forward passes and their backward passes on MFMA per
// 16 samples, with the exact single-electron determinant ratio and drift of
// the Slater-Jastrow wavefunction), the fixed-order gradient reduction and
// Adam.
#include "basis.hpp"
#include "mlp_internal.hpp"

namespace mole {

namespace {

struct LayerDims {
  int woff, nin, nout;
};
__device__ __forceinline__ LayerDims layer_dims(int layer) {
  switch (layer) {
    case 0: return {kMlpW1, kMlpIn, kMlpHid};
    case 1: return {kMlpW2, kMlpHid, kMlpHid};
    case 2: return {kMlpW3, kMlpHid, kMlpHid};
    default: return {kMlpW4, kMlpHid, kMlpOut};
  }
}

__global__ void mlp_init_kernel(float *w, uint64_t seed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  int nin = 0, nout = 0;
  bool weight = true;
  if (p < kMlpB1) { nin = kMlpIn; nout = kMlpHid; }
  else if (p < kMlpW2) weight = false;
  else if (p < kMlpB2) { nin = kMlpHid; nout = kMlpHid; }
  else if (p < kMlpW3) weight = false;
  else if (p < kMlpB3) { nin = kMlpHid; nout = kMlpHid; }
  else if (p < kMlpW4) weight = false;
  else if (p < kMlpB4) { nin = kMlpHid; nout = kMlpOut; }
  else weight = false;
  if (!weight) {
    w[p] = 0.0f;
    return;
  }
  // Xavier uniform: U(-a, a), a = sqrt(6 / (fan_in + fan_out))
  Stream st;
  st.init(seed, 0u, 0u, kMlpInitEpoch);
  const double u = u64_standard(st.draw((uint64_t)p));
  const double a = sqrt(6.0 / (double)(nin + nout));
  w[p] = (float)((2.0 * u - 1.0) * a);
}

// One thread per (fragment, lane) of the forward sets (layer 1 with K = 16,
// layers 2-4 with K = 32 in the order of pack_k32) and the backward set,
// then the biases.
__global__ void mlp_pack_kernel(const float *w, bf16x4 *fwd1, bf16x8 *fwd, bf16x4 *bwd, float *bias) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = kFwd1Frags * 64, nf = kFwdFrags * 64, nb = kBwdFrags * 64;
  if (id < n1) {
    const int mt = id / 64, l = id % 64, r = l % 16, q = l / 16;
    bf16x4 v;
    for (int j = 0; j < 4; ++j) {
      const int m = 16 * mt + r, k = 4 * q + j;  // A[m = out][k = in] = W1[in][out]
      v[j] = k < kMlpIn ? bf16_short(w[kMlpW1 + k * kMlpHid + m]) : (short)0;
    }
    fwd1[id] = v;
  } else if (id < n1 + nf) {
    const int e = id - n1, frag = e / 64, l = e % 64, r = l % 16, q = l / 16;
    int layer, mt, c;
    if (frag < kFwdL3) { layer = 1; mt = frag / 2; c = frag % 2; }
    else if (frag < kFwdL4) { layer = 2; mt = (frag - kFwdL3) / 2; c = (frag - kFwdL3) % 2; }
    else { layer = 3; mt = 0; c = frag - kFwdL4; }
    const LayerDims d = layer_dims(layer);
    bf16x8 v;
    for (int j = 0; j < 8; ++j) {
      const int m = 16 * mt + r;                                          // output feature
      const int k = 32 * c + (j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4));  // input feature
      v[j] = (__bf16)((m < d.nout && k < d.nin) ? w[d.woff + k * d.nout + m] : 0.0f);
    }
    fwd[e] = v;
  } else if (id < n1 + nf + nb) {
    const int e = id - n1 - nf, frag = e / 64, l = e % 64, r = l % 16, q = l / 16;
    int layer, mt, kc;
    if (frag < kBwdL3) { layer = 1; mt = frag / 4; kc = frag % 4; }
    else if (frag < kBwdL4) { layer = 2; mt = (frag - kBwdL3) / 4; kc = (frag - kBwdL3) % 4; }
    else { layer = 3; mt = frag - kBwdL4; kc = 0; }
    const LayerDims d = layer_dims(layer);
    bf16x4 v;
    for (int j = 0; j < 4; ++j) {
      const int m = 16 * mt + r, k = 16 * kc + 4 * q + j;  // A[m = in][k = out] = W[in][out]
      v[j] = (m < d.nin && k < d.nout) ? bf16_short(w[d.woff + m * d.nout + k]) : (short)0;
    }
    bwd[e] = v;
  } else if (id < n1 + nf + nb + kMlpBiasPad) {
    const int b = id - n1 - nf - nb;
    float v = 0.0f;
    if (b < 64) v = w[kMlpB1 + b];
    else if (b < 128) v = w[kMlpB2 + b - 64];
    else if (b < 192) v = w[kMlpB3 + b - 128];
    else if (b - 192 < kMlpOut) v = w[kMlpB4 + b - 192];
    bias[b] = v;
  }
}

// LDS transpose of one accumulator-layout tile T (lane l: T[4q + j][l % 16])
// into the operand layout lane l: T[l % 16][4q + j], rounded to bf16.
__device__ __forceinline__ bf16x4 transpose_tile(f32x4 v, float *stage) {
  const int l = threadIdx.x & 63, r = l % 16, q = l / 16;
  __syncthreads();
  for (int j = 0; j < 4; ++j) stage[(4 * q + j) * 16 + r] = v[j];
  __syncthreads();
  f32x4 t;
  for (int j = 0; j < 4; ++j) t[j] = stage[r * 16 + 4 * q + j];
  return to_bf16x4(t);
}

__device__ __forceinline__ f32x4 bf16x4_to_f32(bf16x4 b) {
  f32x4 r;
  for (int j = 0; j < 4; ++j) r[j] = bf16_float((uint16_t)b[j]);
  return r;
}

// dW[in][out] (+)= Xa^T dZa + Xb^T dZb over the wave's 16 samples of two
// backward passes for one layer (one MFMA chain per tile: pass a, then pass
// b, accumulated in the matrix core), and db[out] (+)= sum_s (dZa + dZb).
// Input tiles X (operand layout, lane l: X^T[in 16 mi + l%16][s 4q+j]) and
// output-gradient tiles D (operand layout of dZ^T).  store: this
// workgroup's first contribution to its partial, written without reading it.
__device__ __forceinline__ void weight_grads2(const bf16x4 *Xa, const bf16x4 *Da, const bf16x4 *Xb,
                                              const bf16x4 *Db, int n_in_tiles, int n_out_tiles, int woff,
                                              int boff, int nin, int nout, float *g, bool store) {
  const int l = threadIdx.x & 63, r = l % 16, q = l / 16;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  bf16x4 ones;
  for (int j = 0; j < 4; ++j) ones[j] = (short)0x3F80;  // bf16 1.0
  for (int no = 0; no < n_out_tiles; ++no) {
    const int out = 16 * no + r;
    for (int mi = 0; mi < n_in_tiles; ++mi) {
      const f32x4 d = mfma16(Xb[mi], Db[no], mfma16(Xa[mi], Da[no], zero));
      for (int j = 0; j < 4; ++j) {
        const int in = 16 * mi + 4 * q + j;
        if (in < nin && out < nout) {
          float *gp = g + woff + in * nout + out;
          *gp = store ? d[j] : *gp + d[j];
        }
      }
    }
    const f32x4 db = mfma16(ones, Db[no], mfma16(ones, Da[no], zero));
    if (q == 0 && out < nout) {
      float *gp = g + boff + out;
      *gp = store ? db[0] : *gp + db[0];
    }
  }
}

// The backward passes of two forward evaluations in lockstep, layer by
// layer: a with output gradient dOa, activation tiles ha (registers) and
// bf16 input operand xa; b with dOb, its activation tiles in LDS (hb[k][l])
// and xb.  Adds (stores, with store) both passes' dO/dtheta into g: each
// weight-gradient element is read and written once for the two passes.
__device__ void mlp_backward2(const MlpDev &net, f32x4 dOa, const f32x4 *ha, bf16x4 xa, f32x4 dOb,
                              const f32x4 (*hb)[64], bf16x4 xb, float *stage, float *g, bool store) {
  const int l = threadIdx.x & 63;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  bf16x4 Xa[4], Da[4], Xb[4], Db[4];
  f32x4 dza[4], dzb[4];
  // layer 4 (64 -> 4)
  Da[0] = transpose_tile(dOa, stage);
  Db[0] = transpose_tile(dOb, stage);
  for (int mi = 0; mi < 4; ++mi) {
    Xa[mi] = transpose_tile(ha[8 + mi], stage);
    Xb[mi] = transpose_tile(hb[8 + mi][l], stage);
  }
  weight_grads2(Xa, Da, Xb, Db, 4, 1, kMlpW4, kMlpB4, kMlpHid, kMlpOut, g, store);
  {
    const bf16x4 oa = to_bf16x4(dOa), ob = to_bf16x4(dOb);
    for (int mt = 0; mt < 4; ++mt) {
      const f32x4 A = mfma16(net.bwd[(kBwdL4 + mt) * 64 + l], oa, zero);
      const f32x4 B = mfma16(net.bwd[(kBwdL4 + mt) * 64 + l], ob, zero);
      const f32x4 tha = ha[8 + mt], thb = hb[8 + mt][l];
      for (int j = 0; j < 4; ++j) {
        dza[mt][j] = A[j] * (1.0f - tha[j] * tha[j]);
        dzb[mt][j] = B[j] * (1.0f - thb[j] * thb[j]);
      }
    }
  }
  // layers 3 and 2 (64 -> 64)
  for (int layer = 2; layer >= 1; --layer) {
    const int hin = layer == 2 ? 4 : 0;  // inputs of this layer: h2 (layer 3) or h1 (layer 2)
    const int woff = layer == 2 ? kMlpW3 : kMlpW2, boff = layer == 2 ? kMlpB3 : kMlpB2;
    const int bfr = layer == 2 ? kBwdL3 : kBwdL2;
    for (int no = 0; no < 4; ++no) {
      Da[no] = transpose_tile(dza[no], stage);
      Db[no] = transpose_tile(dzb[no], stage);
    }
    for (int mi = 0; mi < 4; ++mi) {
      Xa[mi] = transpose_tile(ha[hin + mi], stage);
      Xb[mi] = transpose_tile(hb[hin + mi][l], stage);
    }
    weight_grads2(Xa, Da, Xb, Db, 4, 4, woff, boff, kMlpHid, kMlpHid, g, store);
    bf16x4 za[4], zb[4];
    for (int kc = 0; kc < 4; ++kc) {
      za[kc] = to_bf16x4(dza[kc]);
      zb[kc] = to_bf16x4(dzb[kc]);
    }
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 A = zero, B = zero;
      for (int kc = 0; kc < 4; ++kc) {
        A = mfma16(net.bwd[(bfr + mt * 4 + kc) * 64 + l], za[kc], A);
        B = mfma16(net.bwd[(bfr + mt * 4 + kc) * 64 + l], zb[kc], B);
      }
      const f32x4 tha = ha[hin + mt], thb = hb[hin + mt][l];
      for (int j = 0; j < 4; ++j) {
        dza[mt][j] = A[j] * (1.0f - tha[j] * tha[j]);
        dzb[mt][j] = B[j] * (1.0f - thb[j] * thb[j]);
      }
    }
  }
  // layer 1 (13 -> 64): inputs are the features
  for (int no = 0; no < 4; ++no) {
    Da[no] = transpose_tile(dza[no], stage);
    Db[no] = transpose_tile(dzb[no], stage);
  }
  Xa[0] = transpose_tile(bf16x4_to_f32(xa), stage);
  Xb[0] = transpose_tile(bf16x4_to_f32(xb), stage);
  weight_grads2(Xa, Da, Xb, Db, 1, 4, kMlpW1, kMlpB1, kMlpIn, kMlpHid, g, store);
}

// Spin blocks of the training systems hold at most this many electrons.
constexpr int kMlpMaxSpin = 8;

// Rows of the Slater matrix A of electron i's spin block at cfg and the
// right-hand side e_c of A y = e_c into scr (n^2 + n doubles): the four lanes
// of a sample (sub = 0..3) build rows sub, sub + 4, ... (each row by the
// same operations as one lane would).  A barrier must follow.
__device__ void slater_rows(const SJSysDev &S, const double *cfg, int i, double *scr, int sub) {
  const int sp = i < S.n_up ? 0 : 1;
  const int n = sp == 0 ? S.n_up : S.n_dn, first = sp == 0 ? 0 : S.n_up, c = i - first;
  const int K = S.n_basis, no = S.n_orb;
  double *A = scr, *y = scr + n * n;
  for (int row = sub; row < n; row += 4) {
    const double *re = cfg + 3 * (first + row);
    for (int k = 0; k < n; ++k) A[row * n + k] = 0.0;
    for (int m = 0; m < K; ++m) {
      const int I = S.center[m];
      const double d0 = re[0] - S.nuc[3 * I], d1 = re[1] - S.nuc[3 * I + 1], d2 = re[2] - S.nuc[3 * I + 2];
      const double ch = basis_value<false>(S.any_p ? S.kind[m] : MOLE_BASIS_STO, S.zeta[m], d0, d1, d2,
                                           norm3(d0, d1, d2));
      for (int k = 0; k < n; ++k) A[row * n + k] = fma(S.coef_t[(size_t)m * no + k], ch, A[row * n + k]);
    }
    y[row] = row == c ? 1.0 : 0.0;
  }
}

// ln (psi(x')/psi(x))^2 and grad_{x'} ln|psi(x')| of moving electron i of
// cfg to xn (FP64, one lane), from the rows slater_rows left in scr:
// Gaussian elimination with partial pivoting of A y = e_c gives column c of
// A^-1.
__device__ void move_terms(const SJSysDev &S, const double *cfg, int i, const double *xn, double *scr,
                           double *lR2, double *gx) {
  const int N = S.N, sp = i < S.n_up ? 0 : 1;
  const int n = sp == 0 ? S.n_up : S.n_dn;
  const int K = S.n_basis, no = S.n_orb;
  double *A = scr, *y = scr + n * n;
  for (int p = 0; p < n; ++p) {
    int piv = p;
    for (int r = p + 1; r < n; ++r)
      if (fabs(A[r * n + p]) > fabs(A[piv * n + p])) piv = r;
    if (piv != p) {
      for (int k = 0; k < n; ++k) {
        const double t = A[p * n + k];
        A[p * n + k] = A[piv * n + k];
        A[piv * n + k] = t;
      }
      const double t = y[p];
      y[p] = y[piv];
      y[piv] = t;
    }
    for (int r = p + 1; r < n; ++r) {
      const double f = A[r * n + p] / A[p * n + p];
      for (int k = p; k < n; ++k) A[r * n + k] = fma(-f, A[p * n + k], A[r * n + k]);
      y[r] = fma(-f, y[p], y[r]);
    }
  }
  for (int p = n - 1; p >= 0; --p) {
    double s = y[p];
    for (int k = p + 1; k < n; ++k) s = fma(-A[p * n + k], y[k], s);
    y[p] = s / A[p * n + p];
  }
  // orbitals and their gradients at x': each basis function once, the
  // orbitals' sums over m in m order as before
  double phi[kMlpMaxSpin], g0[kMlpMaxSpin], g1[kMlpMaxSpin], g2[kMlpMaxSpin];
#pragma unroll
  for (int k = 0; k < kMlpMaxSpin; ++k) phi[k] = g0[k] = g1[k] = g2[k] = 0.0;
  for (int m = 0; m < K; ++m) {
    const int I = S.center[m];
    const double d0 = xn[0] - S.nuc[3 * I], d1 = xn[1] - S.nuc[3 * I + 1], d2 = xn[2] - S.nuc[3 * I + 2];
    const double rr = norm3(d0, d1, d2);
    double ch, b0, b1, b2, bl;
    basis_vgl<false>(S.any_p ? S.kind[m] : MOLE_BASIS_STO, S.zeta[m], d0, d1, d2, rr, ch, b0, b1, b2, bl);
    const double *cm = S.coef_t + (size_t)m * no;
#pragma unroll
    for (int k = 0; k < kMlpMaxSpin; ++k)
      if (k < n) {
        phi[k] = fma(cm[k], ch, phi[k]);
        g0[k] = fma(cm[k], b0, g0[k]);
        g1[k] = fma(cm[k], b1, g1[k]);
        g2[k] = fma(cm[k], b2, g2[k]);
      }
  }
  double rd = 0.0, gd[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    rd = fma(phi[k], y[k], rd);
    gd[0] = fma(g0[k], y[k], gd[0]);
    gd[1] = fma(g1[k], y[k], gd[1]);
    gd[2] = fma(g2[k], y[k], gd[2]);
  }
  // Jastrow difference and gradient at x'
  const double *r = cfg + 3 * i;
  const double bA = S.p[0], bP = S.p[1], cc = S.p[2], a = S.p[3], d = S.p[4];
  double dJ = 0.0, gj[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *xj = cfg + 3 * j;
    const double b = ((j < S.n_up ? 0 : 1) == sp) ? bP : bA;
    const double dv0 = xn[0] - xj[0], dv1 = xn[1] - xj[1], dv2 = xn[2] - xj[2];
    const double rn = norm3(dv0, dv1, dv2);
    const double ro = norm3(r[0] - xj[0], r[1] - xj[1], r[2] - xj[2]);
    dJ += (b * rn) / (1.0 + cc * rn) - (b * ro) / (1.0 + cc * ro);
    const double den = 1.0 + cc * rn;
    const double up = b / (den * den);
    gj[0] += up * dv0 / rn;
    gj[1] += up * dv1 / rn;
    gj[2] += up * dv2 / rn;
  }
  for (int I = 0; I < S.n_nuc; ++I) {
    const double aZ = a * S.charge[I];
    const double d0 = xn[0] - S.nuc[3 * I], d1 = xn[1] - S.nuc[3 * I + 1], d2 = xn[2] - S.nuc[3 * I + 2];
    const double rn = norm3(d0, d1, d2);
    const double ro = norm3(r[0] - S.nuc[3 * I], r[1] - S.nuc[3 * I + 1], r[2] - S.nuc[3 * I + 2]);
    dJ += -(aZ * rn) / (1.0 + d * rn) + (aZ * ro) / (1.0 + d * ro);
    const double den = 1.0 + d * rn;
    const double vp = -aZ / (den * den);
    gj[0] += vp * d0 / rn;
    gj[1] += vp * d1 / rn;
    gj[2] += vp * d2 / rn;
  }
  *lR2 = 2.0 * log_m(fabs(rd)) + 2.0 * dJ;
  for (int q = 0; q < 3; ++q) gx[q] = gd[q] / rd + gj[q];
}

// Per-wave LDS of the training kernel.
struct TrainLds {
  static constexpr int kMaxN = 16, kMaxSpin = kMlpMaxSpin;
  double cfg[16][3 * kMaxN];
  double ge[16][kMaxSpin * kMaxSpin + kMaxSpin];
  double smp[16][16];      // per sample: x'[3], z[3], mu[3], ls, out3 ... (see below)
  f32x4 h1[12][64];        // activations of the forward pass at x
  float stage[256];
  double obj[16];
};

__global__ void __launch_bounds__(64) mlp_train_kernel(MlpTrainArgs a) {
  __shared__ TrainLds s;
  const SJSysDev &S = *a.sys;
  const int l = threadIdx.x, sl = l & 15, q = l >> 4;
  const int N = S.N;
  const int n_mb = (int)(a.batch / 16);
  // this workgroup's partial gradient: its first minibatch stores every
  // element (the backward passes cover all kMlpParams), later ones add
  float *g = a.partial + (size_t)blockIdx.x * kMlpParams;
  double obj_acc = 0.0;
  for (int mb = blockIdx.x; mb < n_mb; mb += gridDim.x) {
    const uint64_t gi = a.first + (uint64_t)mb * 16 + sl;
    const int i = (int)(gi % (uint64_t)N);
    const uint32_t c = (uint32_t)((gi / (uint64_t)N) % a.n_cfg);
    __syncthreads();
    for (int k = q; k < 3 * N; k += 4) s.cfg[sl][k] = a.cfg[(size_t)c * 3 * N + k];
    __syncthreads();
    const double *cfg = s.cfg[sl];
    const double *r = cfg + 3 * i;
    // the Slater matrix at x for the determinant ratio (rows over the
    // sample's four lanes; the barrier after the forward pass orders them
    // before move_terms)
    slater_rows(S, cfg, i, s.ge[sl], q);
    // forward at x
    float fv[4];
    mlp_features(cfg, N, S.nuc, S.n_nuc, i, r, q, fv);
    // nearest-nucleus distance at x (feature 0, held by the lanes q == 0)
    const double dn = (double)fv[0];
    const bf16x4 x1 = features_bf16(fv);
    f32x4 h[12];
    const f32x4 o1 = mlp_forward(a.net, x1, h);
    for (int k = 0; k < 12; ++k) s.h1[k][l] = h[k];
    __syncthreads();
    if (q == 0) {
      Stream st;
      st.init(a.seed, kMlpTrainSys, (uint32_t)gi, (uint32_t)(gi >> 32));
      double z[4];
      box_muller(st.draw(0), st.draw(1), &z[0], &z[1]);
      box_muller(st.draw(2), st.draw(3), &z[2], &z[3]);
      const double ls = mlp_log_sigma((double)o1[3]);
      const double sg = exp_m(ls);
      double *sp = s.smp[sl];
      for (int k = 0; k < 3; ++k) {
        sp[k] = (r[k] + (double)o1[k]) + sg * z[k];  // x'
        sp[3 + k] = z[k];
        sp[6 + k] = (double)o1[k] + sg * z[k];       // dx
      }
      sp[9] = ls;
      sp[10] = sg;
      sp[11] = (double)o1[3];
      double lR2, gx[3];
      move_terms(S, cfg, i, sp, s.ge[sl], &lR2, gx);
      sp[12] = lR2;
      sp[13] = gx[0];
      sp[14] = gx[1];
      sp[15] = gx[2];
    }
    __syncthreads();
    // forward at x' (electron i moved; the inputs are constants of the gradient)
    const double *sp = s.smp[sl];
    mlp_features(cfg, N, S.nuc, S.n_nuc, i, sp, q, fv);
    const bf16x4 x2 = features_bf16(fv);
    const f32x4 o2 = mlp_forward(a.net, x2, h);
    f32x4 dO1 = {0.f, 0.f, 0.f, 0.f}, dO2 = {0.f, 0.f, 0.f, 0.f};
    if (q == 0) {
      // the relative objective's weight 1/(d^2 + d0^2), d the larger of the
      // nearest-nucleus distances at x and x' (a jump is measured on the
      // scale of its wider end); a constant of the gradient
      const double dm = dn > (double)fv[0] ? dn : (double)fv[0];
      double wgt = a.objective == MOLE_MLP_OBJ_ESJD_RELATIVE ? 1.0 / (dm * dm + kMlpRelD0 * kMlpRelD0) : 1.0;
      const double ls = sp[9], sg = sp[10], ls2 = mlp_log_sigma((double)o2[3]), sg2 = exp_m(ls2);
      double e[3], dx[3], z[3], gx[3] = {sp[13], sp[14], sp[15]};
      double e2 = 0.0, z2 = 0.0, d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        e[k] = (r[k] - sp[k]) - (double)o2[k];
        dx[k] = sp[6 + k];
        z[k] = sp[3 + k];
        e2 += e[k] * e[k];
        z2 += z[k] * z[k];
        d2 += dx[k] * dx[k];
      }
      const double lq_rev = -3.0 * ls2 - e2 / (2.0 * sg2 * sg2);
      const double lq_fwd = -3.0 * ls - 0.5 * z2;
      const double logr = sp[12] + lq_rev - lq_fwd;
      const double A = logr < 0.0 ? exp_m(logr) : 1.0;
      if (a.objective == MOLE_MLP_OBJ_LOG_ESJD) {
        // d log(eps + O) = dO / (eps + O): the ESJD derivatives scaled by
        // this sample's constant 1 / (eps + O)
        const double f = kMlpLogEps + A * d2;
        wgt = 1.0 / f;
        obj_acc += log_m(f);
      } else {
        obj_acc += wgt * (A * d2);
      }
      // the weight is a constant of the gradient: scale every output's
      // derivative by it
      const double Aact = (logr < 0.0 ? A : 0.0) * wgt;
      const double is2 = 1.0 / (sg2 * sg2);
      double GD[3];
      for (int k = 0; k < 3; ++k) GD[k] = 2.0 * (A * wgt) * dx[k] + d2 * Aact * (2.0 * gx[k] + e[k] * is2);
      const double dls = (GD[0] * z[0] + GD[1] * z[1] + GD[2] * z[2]) * sg + d2 * Aact * 3.0;
      const double o13 = sp[11], o23 = (double)o2[3];
      for (int k = 0; k < 3; ++k) {
        dO1[k] = (float)GD[k];
        dO2[k] = (float)(d2 * Aact * (e[k] * is2));
      }
      dO1[3] = (float)(dls * mlp_log_sigma_slope(o13));
      const double dls2 = d2 * Aact * (-3.0 + e2 * is2);
      dO2[3] = (float)(dls2 * mlp_log_sigma_slope(o23));
    }
    mlp_backward2(a.net, dO2, h, x2, dO1, s.h1, x1, s.stage, g, mb == (int)blockIdx.x);
  }
  // objective: fixed-order sum over the wave's sample lanes
  __syncthreads();
  if (q == 0) s.obj[sl] = obj_acc;
  __syncthreads();
  if (l == 0) {
    double t = 0.0;
    for (int k = 0; k < 16; ++k) t += s.obj[k];
    a.obj[blockIdx.x] = t;
  }
}

// Adam on parameter p from its summed gradient (g = -grad / batch)
__device__ __forceinline__ void adam_update(const MlpAdamStep &a, int p, float grad) {
  const float gr = -grad * a.inv_batch;  // loss = -mean objective
  const float mm = a.beta1 * a.m[p] + (1.0f - a.beta1) * gr;
  const float vv = a.beta2 * a.v[p] + (1.0f - a.beta2) * gr * gr;
  a.m[p] = mm;
  a.v[p] = vv;
  const float mh = mm / (1.0f - powf(a.beta1, (float)a.step));
  const float vh = vv / (1.0f - powf(a.beta2, (float)a.step));
  a.w[p] = a.w[p] - a.lr * mh / (sqrtf(vh) + a.eps);
}

// First stage when many workgroups wrote partials: slice h = blockIdx.y of
// the rows, [h S, min(G, (h + 1) S)) with S = kMlpSlice, summed in row order
// per parameter into mid[h][p] (loads 16 at a time, added in that order),
// and the objective partials of the slice into mid_obj[h] by the thread
// after the last parameter.  Slices in parallel: G = 1,024 partials are
// read by 32 x 38 workgroups instead of 38.
__global__ void mlp_reduce_slices_kernel(const float *partial, const double *obj, int G, float *mid,
                                         double *mid_obj) {
  constexpr int kBatch = 16;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int g0 = blockIdx.y * kMlpSlice, g1 = min(G, g0 + kMlpSlice);
  if (p < kMlpParams) {
    float t = 0.0f;
    int gi = g0;
    for (; gi + kBatch <= g1; gi += kBatch) {
      float v[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) v[q] = partial[(size_t)(gi + q) * kMlpParams + p];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) t += v[q];
    }
    for (; gi < g1; ++gi) t += partial[(size_t)gi * kMlpParams + p];
    mid[(size_t)blockIdx.y * kMlpParams + p] = t;
  } else if (p == kMlpParams) {
    double t = 0.0;
    for (int gi = g0; gi < g1; ++gi) t += obj[gi];
    mid_obj[blockIdx.y] = t;
  }
}

// Per parameter the partials of the G workgroups (or the G slice sums of
// the first stage) summed in order; the loads go out 16 at a time and are
// added in that order.  With do_adam the thread then takes the Adam step
// of its parameter.
__global__ void mlp_reduce_kernel(const float *partial, const double *obj, int G, float *grad, double *obj_sum,
                                  MlpAdamStep adam, int do_adam) {
  constexpr int kBatch = 16;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) {
    float t = 0.0f;
    int gi = 0;
    for (; gi + kBatch <= G; gi += kBatch) {
      float v[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) v[q] = partial[(size_t)(gi + q) * kMlpParams + p];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) t += v[q];
    }
    for (; gi < G; ++gi) t += partial[(size_t)gi * kMlpParams + p];
    grad[p] = t;
    if (do_adam) adam_update(adam, p, t);
  }
  if (p == 0) {
    double t = 0.0;
    int gi = 0;
    for (; gi + kBatch <= G; gi += kBatch) {
      double v[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) v[q] = obj[gi + q];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) t += v[q];
    }
    for (; gi < G; ++gi) t += obj[gi];
    *obj_sum = t;
  }
}


// out[4 e + k] = network output k for electron electron[e] of configuration e
__global__ void __launch_bounds__(64) mlp_eval_kernel(MlpDev net, const double *cfg, int N, const double *nuc,
                                                      int n_nuc, const int *electron, int n, float *out) {
  const int l = threadIdx.x, sl = l & 15, q = l >> 4;
  const int e = blockIdx.x * 16 + sl;
  float fv[4] = {0.f, 0.f, 0.f, 0.f};
  if (e < n) {
    const double *c = cfg + (size_t)e * 3 * N;
    const int i = electron[e];
    mlp_features(c, N, nuc, n_nuc, i, c + 3 * i, q, fv);
  }
  const f32x4 o = mlp_forward(net, features_bf16(fv), nullptr);
  if (q == 0 && e < n)
    for (int k = 0; k < 4; ++k) out[4 * e + k] = o[k];
}

}  // namespace

hipError_t launch_mlp_init(float *w, uint64_t seed, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_init_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, w, seed);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float *w, bf16x4 *fwd1, bf16x8 *fwd, bf16x4 *bwd, float *bias,
                           hipStream_t stream) {
  const int n = (kFwd1Frags + kFwdFrags + kBwdFrags) * 64 + kMlpBiasPad;
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, w, fwd1, fwd, bwd, bias);
  return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpTrainArgs &a, int n_groups, hipStream_t stream) {
  hipLaunchKernelGGL(mlp_train_kernel, dim3(n_groups), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce(const float *partial, const double *obj, int n_groups, float *mid, double *mid_obj,
                             float *grad, double *obj_sum, const MlpAdamStep *adam, hipStream_t stream) {
  const MlpAdamStep none{};
  if (n_groups > kMlpSlice) {
    const int H = (n_groups + kMlpSlice - 1) / kMlpSlice;
    hipLaunchKernelGGL(mlp_reduce_slices_kernel, dim3((kMlpParams + 1 + 255) / 256, H), dim3(256), 0, stream,
                       partial, obj, n_groups, mid, mid_obj);
    partial = mid;
    obj = mid_obj;
    n_groups = H;
  }
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, partial, obj,
                     n_groups, grad, obj_sum, adam ? *adam : none, adam ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_mlp_eval(const MlpDev &net, const double *cfg, int N, const double *nuc, int n_nuc,
                           const int *electron, int n, float *out, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_eval_kernel, dim3((n + 15) / 16), dim3(64), 0, stream, net, cfg, N, nuc, n_nuc,
                     electron, n, out);
  return hipGetLastError();
}

}  // namespace mole
