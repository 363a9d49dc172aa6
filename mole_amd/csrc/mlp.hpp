// Walker MLP of the learned Metropolis-Hastings proposal: 13 -> 64 -> 64 ->
// 64 -> 4 with tanh hidden layers, bf16 operands and FP32 accumulation on
// v_mfma_f32_16x16x16_bf16.
//
// One wave evaluates 16 samples.  Every layer is computed transposed,
// Z^T = W^T H^T (M = output features, N = samples, K = input features):
// the accumulator of a 16-feature output tile (lane l holds features
// 4(l/16)..+3 of sample l%16) is, after tanh and bf16 rounding, exactly the
// B operand of the next layer's K-chunk, so no data moves between layers.
// Sample s of the wave lives in lanes s, s+16, s+32, s+48 (lane l: sample
// l%16, sub-lane l/16), which also hold its input features 4(l/16)..+3.
//
// Weights: FP32 master copy in the reference order W1[13][64] b1 W2[64][64]
// b2 W3[64][64] b3 W4[64][4] b4; the kernels read bf16 fragments packed from
// it (one 8-byte load per lane per MFMA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace mole {

constexpr int kMlpIn = 13, kMlpHid = 64, kMlpOut = 4;
constexpr int kMlpParams = 9476;
constexpr int kMlpW1 = 0, kMlpB1 = 832, kMlpW2 = 896, kMlpB2 = 4992, kMlpW3 = 5056, kMlpB3 = 9152,
              kMlpW4 = 9216, kMlpB4 = 9472;
constexpr double kMlpLS0 = -1.3862943611198906;  // ln 0.25
constexpr double kMlpLSHalf = 3.0;                // ln sigma within ln 0.25 +- 3
// forward fragments: layer 1 on 16x16x16 [4 m-tiles]; layers 2-4 on
// 16x16x32: layers 2-3 [4 m][2 k], layer 4 [2 k]
constexpr int kFwd1Frags = 4;
constexpr int kFwdL2 = 0, kFwdL3 = 8, kFwdL4 = 16, kFwdFrags = 18;
// backward fragments (A = W, M = inputs): layers 2-3 [4 m][4 k], layer 4 [4 m]
constexpr int kBwdL2 = 0, kBwdL3 = 16, kBwdL4 = 32, kBwdFrags = 36;
constexpr int kMlpBiasPad = 64 + 64 + 64 + 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MlpDev {
  const bf16x4 *fwd1;  // [kFwd1Frags][64]   layer 1, K = 16
  const bf16x8 *fwd;   // [kFwdFrags][64]    layers 2-4, K = 32
  const bf16x4 *bwd;   // [kBwdFrags][64]
  const float *bias;   // b1[64] b2[64] b3[64] b4[16] (zero padded)
};

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16) + j] and
// B[8(l/16) + j][l%16], j = 0..7; the accumulator as for 16x16x16
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short bf16_short(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__device__ __forceinline__ bf16x4 to_bf16x4(f32x4 v) {
  bf16x4 r;
  r[0] = bf16_short(v[0]);
  r[1] = bf16_short(v[1]);
  r[2] = bf16_short(v[2]);
  r[3] = bf16_short(v[3]);
  return r;
}

// Two accumulator tiles of consecutive 16-feature blocks as one K = 32
// operand: position 8q + j holds feature 4q + j of the first tile (j < 4)
// or 4q + j - 4 of the second; the K = 32 weight fragments are packed in
// this order.
__device__ __forceinline__ bf16x8 pack_k32(f32x4 lo, f32x4 hi) {
  bf16x8 r;
  r[0] = (__bf16)lo[0];
  r[1] = (__bf16)lo[1];
  r[2] = (__bf16)lo[2];
  r[3] = (__bf16)lo[3];
  r[4] = (__bf16)hi[0];
  r[5] = (__bf16)hi[1];
  r[6] = (__bf16)hi[2];
  r[7] = (__bf16)hi[3];
  return r;
}

// tanh x = 1 - 2 / (e^{2x} + 1) on the hardware exp2 and reciprocal
// (absolute error ~2e-7, far below the bf16 rounding of the activations);
// saturates to +-1 for large |x|.
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);  // 2 log2(e)
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ __forceinline__ f32x4 tanh4(f32x4 v) {
  f32x4 r;
  r[0] = tanh_fast(v[0]);
  r[1] = tanh_fast(v[1]);
  r[2] = tanh_fast(v[2]);
  r[3] = tanh_fast(v[3]);
  return r;
}

__device__ __forceinline__ f32x4 bias4(const float *b, int row0) {
  return *(const f32x4 *)(b + row0);
}

// Forward pass of the wave's 16 samples.  x: this lane's four input
// features (bf16).  h (optional): the hidden activations after tanh, [3][4]
// tiles in accumulator layout.  Returns the output tile: lanes 0..15 hold
// out[0..3] of samples 0..15.  All 64 lanes must call it.
__device__ __forceinline__ f32x4 mlp_forward(const MlpDev &m, bf16x4 x, f32x4 *h) {
  const int l = threadIdx.x & 63, q = l >> 4;
  // every weight fragment of the pass is loaded before the first MFMA, so
  // the loads are in flight together instead of one layer at a time
  bf16x4 w1[4];
  bf16x8 wf[kFwdFrags];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) w1[mt] = m.fwd1[mt * 64 + l];
#pragma unroll
  for (int f = 0; f < kFwdFrags; ++f) wf[f] = m.fwd[f * 64 + l];
  f32x4 t[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = bias4(m.bias, 16 * mt + 4 * q);
    acc = mfma16(w1[mt], x, acc);
    t[mt] = tanh4(acc);
    if (h) h[mt] = t[mt];
  }
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const bf16x8 b0 = pack_k32(t[0], t[1]), b1 = pack_k32(t[2], t[3]);
    const int fr = layer == 0 ? kFwdL2 : kFwdL3;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = bias4(m.bias + 64 * (layer + 1), 16 * mt + 4 * q);
      acc = mfma32(wf[fr + mt * 2], b0, acc);
      acc = mfma32(wf[fr + mt * 2 + 1], b1, acc);
      t[mt] = tanh4(acc);
      if (h) h[4 * (layer + 1) + mt] = t[mt];
    }
  }
  f32x4 out = bias4(m.bias + 192, 4 * q);
  out = mfma32(wf[kFwdL4], pack_k32(t[0], t[1]), out);
  out = mfma32(wf[kFwdL4 + 1], pack_k32(t[2], t[3]), out);
  return out;
}

// tanh(out3 / 3) as 1 - 2 / (e^{2y} + 1) on exp_m (absolute error ~1e-16;
// +-1 beyond |y| = 20)
__device__ __forceinline__ double mlp_sigma_tanh(double out3) {
  const double y = fmin(fmax(out3 / kMlpLSHalf, -20.0), 20.0);
  return 1.0 - 2.0 / (exp_m(2.0 * y) + 1.0);
}

// ln sigma from the raw output: ln 0.25 + 3 tanh(out3 / 3), a smooth bound
// (sigma in [0.0124, 5.02]) whose slope never vanishes, so a proposal
// pushed towards either end can still be trained back
__device__ __forceinline__ double mlp_log_sigma(double out3) {
  return kMlpLS0 + kMlpLSHalf * mlp_sigma_tanh(out3);
}

// d ln sigma / d out3
__device__ __forceinline__ double mlp_log_sigma_slope(double out3) {
  const double t = mlp_sigma_tanh(out3);
  return 1.0 - t * t;
}

// The 13 inputs of electron i at r (FP64; rounded to bf16 by the caller):
// nearest-nucleus distance, x_j - r for the three nearest other electrons
// (nearest first, ties to the lower index, zeros when missing), r.
// Writes features 4 sub .. 4 sub + 3 into f (selected without a runtime
// index into a private array, which would live in scratch memory).
__device__ __forceinline__ void mlp_features(const double *pos, int N, const double *nuc, int n_nuc, int i,
                                             const double *r, int sub, float *f) {
  const double kInf = __longlong_as_double(0x7FF0000000000000ll);
  double dmin = kInf;
  for (int I = 0; I < n_nuc; ++I) {
    const double d = norm3(r[0] - nuc[3 * I], r[1] - nuc[3 * I + 1], r[2] - nuc[3 * I + 2]);
    dmin = d < dmin ? d : dmin;
  }
  int b0 = -1, b1 = -1, b2 = -1;
  double d0 = kInf, d1 = kInf, d2 = kInf;
  // insertion into the sorted three by selects (a later equal distance does
  // not displace an earlier one: ties go to the lower index)
  auto insert = [&](double d, int j) {
    const bool c0 = d < d0, c1 = d < d1, c2 = d < d2;
    d2 = c1 ? d1 : (c2 ? d : d2);
    b2 = c1 ? b1 : (c2 ? j : b2);
    d1 = c0 ? d0 : (c1 ? d : d1);
    b1 = c0 ? b0 : (c1 ? j : b1);
    d0 = c0 ? d : d0;
    b0 = c0 ? j : b0;
  };
  // squared distance (same order as the distance, no square root); ties of
  // the squared values are ties of the distances.  Positions are loaded four
  // partners at a time ahead of the arithmetic (the scan order is unchanged).
  auto dist2 = [&](const double *x) {
    const double e0 = x[0] - r[0], e1 = x[1] - r[1], e2 = x[2] - r[2];
    double d = e0 * e0;
    d = d + e1 * e1;
    d = d + e2 * e2;
    return d;
  };
  int j = 0;
  for (; j + 4 <= N; j += 4) {
    double xv[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 3; ++q) xv[u][q] = pos[3 * (j + u) + q];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double d = dist2(xv[u]);
      // the moved electron itself never enters
      insert(j + u == i ? kInf : d, j + u == i ? -1 : j + u);
    }
  }
  for (; j < N; ++j)
    if (j != i) insert(dist2(pos + 3 * j), j);
  auto vec = [&](int b, int q) { return b >= 0 ? pos[3 * b + q] - r[q] : 0.0; };
  double v0, v1, v2, v3;
  if (sub == 0) {
    v0 = dmin; v1 = vec(b0, 0); v2 = vec(b0, 1); v3 = vec(b0, 2);
  } else if (sub == 1) {
    v0 = vec(b1, 0); v1 = vec(b1, 1); v2 = vec(b1, 2); v3 = vec(b2, 0);
  } else if (sub == 2) {
    v0 = vec(b2, 1); v1 = vec(b2, 2); v2 = r[0]; v3 = r[1];
  } else {
    v0 = r[2]; v1 = 0.0; v2 = 0.0; v3 = 0.0;
  }
  f[0] = (float)v0;
  f[1] = (float)v1;
  f[2] = (float)v2;
  f[3] = (float)v3;
}

__device__ __forceinline__ bf16x4 features_bf16(const float *f) {
  bf16x4 r;
  for (int j = 0; j < 4; ++j) r[j] = bf16_short(f[j]);
  return r;
}

}  // namespace mole
