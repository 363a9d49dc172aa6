// Diagnostic kernels: the device primitives of the walker path applied to
// caller-given inputs, for element-wise comparison with the CPU restatement.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/mole_gpu_debug.h"
#include "device_common.hpp"
#include "mlp.hpp"

using namespace mole;

namespace {

__global__ void philox_kernel(uint64_t seed, uint32_t sys, uint32_t walker, uint32_t epoch,
                              uint64_t first, size_t count, uint64_t *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  Stream s;
  s.init(seed, sys, walker, epoch);
  out[i] = s.draw(first + i);
}

__global__ void math_kernel(int op, size_t n, const double *a, const double *b, double *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i];
  const double y = b ? b[i] : 0.0;
  double r = 0.0;
  switch (op) {
    case MOLE_DEBUG_EXP: r = exp_m(x); break;
    case MOLE_DEBUG_SQRT: r = sqrt(x); break;
    case MOLE_DEBUG_DIV: r = x / y; break;
    case MOLE_DEBUG_STANDARD: r = u64_standard((uint64_t)__double_as_longlong(x)); break;
    case MOLE_DEBUG_RANGE: {
      const double lo = -y, hi = y;
      const double scale = hi - lo;
      r = u64_range((uint64_t)__double_as_longlong(x), scale, lo - scale);
      break;
    }
    case MOLE_DEBUG_NORM3: r = norm3(x, y, x * y); break;
    case MOLE_DEBUG_LOG: r = log_m(x); break;
    case MOLE_DEBUG_SINPI:
    case MOLE_DEBUG_COSPI: {
      double sn, cs;
      sincospi_m(x, &sn, &cs);
      r = op == MOLE_DEBUG_SINPI ? sn : cs;
      break;
    }
    case MOLE_DEBUG_BOXMULLER0:
    case MOLE_DEBUG_BOXMULLER1: {
      double z0, z1;
      box_muller((uint64_t)__double_as_longlong(x), (uint64_t)__double_as_longlong(y), &z0, &z1);
      r = op == MOLE_DEBUG_BOXMULLER0 ? z0 : z1;
      break;
    }
    case MOLE_DEBUG_BF16: r = (double)bf16_float(bf16_bits((float)x)); break;
    default: r = 0.0;
  }
  out[i] = r;
}

// One v_mfma_f32_16x16x16_bf16 on 16x16 operands given row-major in FP32
// (rounded to bf16 here), placed in lanes as the walker MLP assumes:
//   A: lane l holds A[l % 16][4 (l / 16) + j], j = 0..3
//   B: lane l holds B[4 (l / 16) + j][l % 16]
//   D: lane l holds D[4 (l / 16) + j][l % 16]
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(64) mfma_probe_kernel(const float *A, const float *B, float *D) {
  const int l = threadIdx.x, r = l % 16, q = l / 16;
  bf16x4_t a, b;
  for (int j = 0; j < 4; ++j) {
    a[j] = (short)bf16_bits(A[r * 16 + 4 * q + j]);
    b[j] = (short)bf16_bits(B[(4 * q + j) * 16 + r]);
  }
  const f32x4_t d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, (f32x4_t){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int j = 0; j < 4; ++j) D[(4 * q + j) * 16 + r] = d[j];
}

// One v_mfma_f32_16x16x32_bf16: A 16x32, B 32x16 row-major FP32 (rounded
// to bf16), lane l holding A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16],
// j = 0..7; D as for 16x16x16.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
__global__ void __launch_bounds__(64) mfma32_probe_kernel(const float *A, const float *B, float *D) {
  const int l = threadIdx.x, r = l % 16, q = l / 16;
  bf16x8_t a, b;
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)A[r * 32 + 8 * q + j];
    b[j] = (__bf16)B[(8 * q + j) * 16 + r];
  }
  const f32x4_t d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, (f32x4_t){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int j = 0; j < 4; ++j) D[(4 * q + j) * 16 + r] = d[j];
}

// Plain copy of n doubles: the calibration kernel for HBM byte counters.
__global__ void copy_probe_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                  size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// mlp_features of n (configuration, electron) samples, 16 floats per
// sample: lane sub of the sample's four writes features 4 sub .. 4 sub + 3
__global__ void __launch_bounds__(64) mlp_features_probe_kernel(const double *cfg, int N, const double *nuc,
                                                                int n_nuc, const int *electron, int n,
                                                                float *out) {
  const int l = threadIdx.x, sl = l & 15, q = l >> 4;
  const int e = blockIdx.x * 16 + sl;
  if (e < n) {
    const double *c = cfg + (size_t)e * 3 * N;
    float fv[4];
    mlp_features(c, N, nuc, n_nuc, electron[e], c + 3 * electron[e], q, fv);
    for (int k = 0; k < 4; ++k) out[16 * (size_t)e + 4 * q + k] = fv[k];
  }
}

__global__ void iota_kernel(double *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = (double)i;
}

}  // namespace

extern "C" mole_status mole_debug_copy(size_t n_doubles, int32_t repeats, double *ms_per_copy) {
  if (n_doubles == 0 || repeats < 1) return MOLE_ERR_INVALID_ARGUMENT;
  double *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n_doubles * sizeof(double)) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  if (hipMalloc(&dst, n_doubles * sizeof(double)) != hipSuccess) {
    (void)hipFree(src);
    return MOLE_ERR_OUT_OF_MEMORY;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(iota_kernel, dim3(2048), dim3(256), 0, 0, src, n_doubles);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < repeats; ++r)
    hipLaunchKernelGGL(copy_probe_kernel, dim3(2048), dim3(256), 0, 0, src, dst, n_doubles);
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (ms_per_copy) *ms_per_copy = (double)ms / repeats;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}

extern "C" mole_status mole_debug_philox(uint64_t seed, uint32_t sys, uint32_t walker,
                                         uint32_t epoch, uint64_t first, size_t count,
                                         uint64_t *out) {
  if (!out) return MOLE_ERR_INVALID_ARGUMENT;
  if (count == 0) return MOLE_OK;
  uint64_t *d = nullptr;
  if (hipMalloc(&d, count * sizeof(uint64_t)) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  hipLaunchKernelGGL(philox_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, 0, seed,
                     sys, walker, epoch, first, count, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, count * sizeof(uint64_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}

extern "C" mole_status mole_debug_math(int32_t op, size_t n, const double *a, const double *b,
                                       double *out) {
  if (!a || !out || op < 0 || op > MOLE_DEBUG_BF16) return MOLE_ERR_INVALID_ARGUMENT;
  if ((op == MOLE_DEBUG_DIV || op == MOLE_DEBUG_RANGE || op == MOLE_DEBUG_NORM3 ||
       op == MOLE_DEBUG_BOXMULLER0 || op == MOLE_DEBUG_BOXMULLER1) && !b)
    return MOLE_ERR_INVALID_ARGUMENT;
  if (n == 0) return MOLE_OK;
  double *d = nullptr;
  if (hipMalloc(&d, 3 * n * sizeof(double)) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  hipError_t e = hipMemcpy(d, a, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && b) e = hipMemcpy(d + n, b, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op, n, d,
                       b ? d + n : nullptr, d + 2 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}

extern "C" mole_status mole_debug_mfma_bf16(const float *A, const float *B, float *D) {
  if (!A || !B || !D) return MOLE_ERR_INVALID_ARGUMENT;
  float *d = nullptr;
  if (hipMalloc(&d, 3 * 256 * sizeof(float)) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  hipError_t e = hipMemcpy(d, A, 256 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 256, B, 256 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, 0, d, d + 256, d + 512);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(D, d + 512, 256 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}

extern "C" mole_status mole_debug_mfma_bf16_k32(const float *A, const float *B, float *D) {
  if (!A || !B || !D) return MOLE_ERR_INVALID_ARGUMENT;
  float *d = nullptr;
  if (hipMalloc(&d, (512 + 512 + 256) * sizeof(float)) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  hipError_t e = hipMemcpy(d, A, 512 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 512, B, 512 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mfma32_probe_kernel, dim3(1), dim3(64), 0, 0, d, d + 512, d + 1024);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(D, d + 1024, 256 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}

extern "C" mole_status mole_debug_mlp_features(const double *cfg, uint32_t N, const double *nuc_xyz,
                                               uint32_t n_nuc, const int32_t *electron, size_t n, float *out) {
  if (!cfg || !electron || !out || (n_nuc && !nuc_xyz)) return MOLE_ERR_INVALID_ARGUMENT;
  if (N < 1 || N > MOLE_MAX_ELECTRONS || n_nuc > MOLE_MAX_NUCLEI || n > ((size_t)1 << 24)) return MOLE_ERR_SHAPE;
  for (size_t e = 0; e < n; ++e)
    if (electron[e] < 0 || electron[e] >= (int32_t)N) return MOLE_ERR_INVALID_ARGUMENT;
  if (n == 0) return MOLE_OK;
  const size_t nc = n * 3 * N, nn = 3 * (size_t)n_nuc;
  void *d = nullptr;
  const size_t bytes = (nc + nn) * sizeof(double) + n * sizeof(int32_t) + 16 * n * sizeof(float);
  if (hipMalloc(&d, bytes) != hipSuccess) return MOLE_ERR_OUT_OF_MEMORY;
  double *dc = (double *)d, *dn = dc + nc;
  int32_t *de = (int32_t *)(dn + nn);
  float *dout = (float *)(de + n);
  hipError_t e = hipMemcpy(dc, cfg, nc * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && nn) e = hipMemcpy(dn, nuc_xyz, nn * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(de, electron, n * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp_features_probe_kernel, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, 0, dc, (int)N, dn,
                       (int)n_nuc, de, (int)n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, 16 * n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? MOLE_OK : MOLE_ERR_DEVICE;
}
