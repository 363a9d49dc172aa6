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
constexpr double kMlpLSLo = -7.0, kMlpLSHi = 1.0;
// forward fragments: layer 1 [4 m-tiles], layers 2-3 [4 m][4 k], layer 4 [4 k]
constexpr int kFwdL1 = 0, kFwdL2 = 4, kFwdL3 = 20, kFwdL4 = 36, kFwdFrags = 40;
// backward fragments (A = W, M = inputs): layers 2-3 [4 m][4 k], layer 4 [4 m]
constexpr int kBwdL2 = 0, kBwdL3 = 16, kBwdL4 = 32, kBwdFrags = 36;
constexpr int kMlpBiasPad = 64 + 64 + 64 + 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MlpDev {
  const bf16x4 *fwd;  // [kFwdFrags][64]
  const bf16x4 *bwd;  // [kBwdFrags][64]
  const float *bias;  // b1[64] b2[64] b3[64] b4[16] (zero padded)
};

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 to_bf16x4(f32x4 v) {
  bf16x4 r;
  r[0] = (short)bf16_bits(v[0]);
  r[1] = (short)bf16_bits(v[1]);
  r[2] = (short)bf16_bits(v[2]);
  r[3] = (short)bf16_bits(v[3]);
  return r;
}

__device__ __forceinline__ f32x4 tanh4(f32x4 v) {
  f32x4 r;
  r[0] = tanhf(v[0]);
  r[1] = tanhf(v[1]);
  r[2] = tanhf(v[2]);
  r[3] = tanhf(v[3]);
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
  bf16x4 a[4], b[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = bias4(m.bias, 16 * mt + 4 * q);
    acc = mfma16(m.fwd[(kFwdL1 + mt) * 64 + l], x, acc);
    const f32x4 t = tanh4(acc);
    if (h) h[mt] = t;
    a[mt] = to_bf16x4(t);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = bias4(m.bias + 64, 16 * mt + 4 * q);
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc = mfma16(m.fwd[(kFwdL2 + mt * 4 + kc) * 64 + l], a[kc], acc);
    const f32x4 t = tanh4(acc);
    if (h) h[4 + mt] = t;
    b[mt] = to_bf16x4(t);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = bias4(m.bias + 128, 16 * mt + 4 * q);
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc = mfma16(m.fwd[(kFwdL3 + mt * 4 + kc) * 64 + l], b[kc], acc);
    const f32x4 t = tanh4(acc);
    if (h) h[8 + mt] = t;
    a[mt] = to_bf16x4(t);
  }
  f32x4 out = bias4(m.bias + 192, 4 * q);
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) out = mfma16(m.fwd[(kFwdL4 + kc) * 64 + l], a[kc], out);
  return out;
}

// ln sigma from the raw output (clamped)
__device__ __forceinline__ double mlp_log_sigma(double out3) {
  const double c = out3 < kMlpLSLo ? kMlpLSLo : (out3 > kMlpLSHi ? kMlpLSHi : out3);
  return kMlpLS0 + c;
}

// The 13 inputs of electron i at r (FP64; rounded to bf16 by the caller):
// nearest-nucleus distance, x_j - r for the three nearest other electrons
// (nearest first, ties to the lower index, zeros when missing), r.
// Writes features 4 sub .. 4 sub + 3 into f.
__device__ __forceinline__ void mlp_features(const double *pos, int N, const double *nuc, int n_nuc, int i,
                                             const double *r, int sub, float *f) {
  double v[4];
  if (sub == 0) {
    double dmin = __longlong_as_double(0x7FF0000000000000ll);
    for (int I = 0; I < n_nuc; ++I) {
      const double d = norm3(r[0] - nuc[3 * I], r[1] - nuc[3 * I + 1], r[2] - nuc[3 * I + 2]);
      dmin = d < dmin ? d : dmin;
    }
    v[0] = dmin;
  }
  int best[3] = {-1, -1, -1};
  double bd[3];
  bd[0] = bd[1] = bd[2] = __longlong_as_double(0x7FF0000000000000ll);
  for (int j = 0; j < N; ++j) {
    if (j == i) continue;
    const double *x = pos + 3 * j;
    const double d = norm3(x[0] - r[0], x[1] - r[1], x[2] - r[2]);
    if (d < bd[0]) {
      bd[2] = bd[1]; best[2] = best[1];
      bd[1] = bd[0]; best[1] = best[0];
      bd[0] = d; best[0] = j;
    } else if (d < bd[1]) {
      bd[2] = bd[1]; best[2] = best[1];
      bd[1] = d; best[1] = j;
    } else if (d < bd[2]) {
      bd[2] = d; best[2] = j;
    }
  }
  double all[13];
  all[0] = sub == 0 ? v[0] : 0.0;
  for (int k = 0; k < 3; ++k)
    for (int q = 0; q < 3; ++q) all[1 + 3 * k + q] = best[k] >= 0 ? pos[3 * best[k] + q] - r[q] : 0.0;
  for (int q = 0; q < 3; ++q) all[10 + q] = r[q];
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * sub + j;
    f[j] = k < 13 ? (float)all[k] : 0.0f;
  }
}

__device__ __forceinline__ bf16x4 features_bf16(const float *f) {
  bf16x4 r;
  for (int j = 0; j < 4; ++j) r[j] = (short)bf16_bits(f[j]);
  return r;
}

}  // namespace mole
