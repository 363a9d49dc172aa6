// Integrated autocorrelation time of every walker's local-energy series on
// the device: the estimator of scripts/statfor.rs:30-54,
//   tau = max(1, 1 + 2 sum_{i=1..min(200, n-1)} rho_i f_i),
//   rho_i = sum_{t<n-i} (x_t - m)(x_{t+i} - m) / (var (n - i)),
//   f_i = 0 from the first negative rho on, var with n - 1 (statfor.rs:86-88),
// so effective sample counts need no copy of the series to the host.
//
// One 256-thread workgroup per walker.  The series is staged in LDS; the
// mean and the variance are sums in sample order on thread 0; thread i - 1
// forms rho_i as one sum in sample order (a lag per thread, every sum the
// reference's sequential fold), and thread 0 then adds them in lag order.
// Compiled with -ffp-contract=off, the result equals the CPU fold bit for
// bit.
#include <hip/hip_runtime.h>

#include "../../include/mole_gpu.h"
#include "internal.hpp"

namespace mole {

namespace {

constexpr int kTauThreads = 256;
constexpr int kTauMaxLag = 200;
constexpr int kTauLdsSamples = 8192;  // series up to this length are staged in LDS

template <bool LDS>
__global__ void __launch_bounds__(kTauThreads) tau_corr_kernel(const double *trace_e, size_t cap, size_t first, int n,
                                                               double *tau_out) {
  extern __shared__ double xs[];
  __shared__ double corr[kTauMaxLag];
  __shared__ double mv[2];
  const size_t w = blockIdx.x;
  const double *src = trace_e + w * cap + first;
  const double *x = src;
  if constexpr (LDS) {
    for (int t = threadIdx.x; t < n; t += kTauThreads) xs[t] = src[t];
    __syncthreads();
    x = xs;
  }
  if (threadIdx.x == 0) {
    double mean = 0.0;
    for (int t = 0; t < n; ++t) mean += x[t];
    mean /= (double)n;
    double var = 0.0;
    for (int t = 0; t < n; ++t) var += (mean - x[t]) * (mean - x[t]);
    var /= (double)(n - 1);
    mv[0] = mean;
    mv[1] = var;
  }
  __syncthreads();
  const double mean = mv[0], var = mv[1];
  const int max_i = kTauMaxLag < n - 1 ? kTauMaxLag : n - 1;
  const int i = threadIdx.x + 1;
  if (i <= max_i) {
    double c = 0.0;
    const int m = n - i;
    int t = 0;
    // loads issued eight at a time ahead of the additions (same order)
    for (; t + 8 <= m; t += 8) {
      double a[8], b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        a[u] = x[t + u];
        b[u] = x[t + u + i];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) c += (a[u] - mean) * (b[u] - mean);
    }
    for (; t < m; ++t) c += (x[t] - mean) * (x[t + i] - mean);
    corr[i - 1] = c / (var * (double)m);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tau = 1.0, f = 1.0;
    for (int k = 1; k <= max_i; ++k) {
      const double c = corr[k - 1];
      if (c < 0.0) f = 0.0;
      tau += 2.0 * c * f;
    }
    tau_out[w] = tau > 1.0 ? tau : 1.0;
  }
}

}  // namespace

hipError_t launch_tau_corr(const double *trace_e, int W, size_t cap, size_t first, int n, double *tau_out,
                           hipStream_t stream) {
  if (W <= 0) return hipSuccess;
  if (n <= kTauLdsSamples)
    hipLaunchKernelGGL(tau_corr_kernel<true>, dim3(W), dim3(kTauThreads), sizeof(double) * n, stream, trace_e, cap,
                       first, n, tau_out);
  else
    hipLaunchKernelGGL(tau_corr_kernel<false>, dim3(W), dim3(kTauThreads), 0, stream, trace_e, cap, first, n,
                       tau_out);
  return hipGetLastError();
}

}  // namespace mole
