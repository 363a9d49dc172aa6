// Diffusion Monte Carlo around the Slater-Jastrow DMC steps: the ensemble
// reductions, SRBrancher and SimpleBranching (src/dmc/src/branching.rs:7-91),
// the gather of walker state after branching, and update_energies
// (src/dmc/src/dmc.rs:155-202).  Every kernel reads its counts from the
// device state, so one step's launches can be captured and replayed.
//
// Sums run in a fixed order (per-thread strided sums, then an LDS tree), so
// results do not depend on scheduling.  A uniform integer in [0, n) from a
// 64-bit draw d is the high word of d n (no rejection loop: bias below
// n / 2^64).
#include "device_common.hpp"
#include "dmc_internal.hpp"

namespace mole {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedChunk = kRedThreads * 16;  // walkers per reduction block
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanBlock = kScanThreads * kScanPerThread;

__device__ __forceinline__ uint64_t uniform_below(uint64_t d, uint64_t n) { return __umul64hi(d, n); }

// Rust's `f64 as u32` / `f64 as usize`: truncation, saturating, NaN -> 0
__device__ __forceinline__ uint64_t as_u32(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 4294967295.0) return 0xFFFFFFFFull;
  return (uint64_t)(uint32_t)x;
}
__device__ __forceinline__ uint64_t as_usize(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 18446744073709551615.0) return ~0ull;
  return (uint64_t)x;
}

__global__ void init_configs_kernel(double *pos, int N, int W, uint64_t seed, uint32_t sys) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  Stream st;
  st.init(seed, sys, 0xFFFFFFFFu, kDmcInitEpoch);
  for (int k = 0; 2 * k < 3 * N; ++k) {
    double z0, z1;
    box_muller(st.draw(2 * (uint64_t)k), st.draw(2 * (uint64_t)k + 1), &z0, &z1);
    pos[(size_t)(2 * k) * W + w] = z0;
    if (2 * k + 1 < 3 * N) pos[(size_t)(2 * k + 1) * W + w] = z1;
  }
}

// ---- reductions: sum w E_old, sum w_old, sum w_new, max w_new ----

__global__ void __launch_bounds__(kRedThreads) reduce1_kernel(const double *dmc, int W, double *part, int n_parts) {
  __shared__ double sh[4][kRedThreads];
  const int t = threadIdx.x;
  const size_t b0 = (size_t)blockIdx.x * kRedChunk;
  // every load is issued before the first sum (an index past W reads
  // element 0 and is not summed), so the 16 rounds of a thread overlap in
  // memory instead of waiting one after another; the sums keep their order
  constexpr int kPer = kRedChunk / kRedThreads;
  double vwe[kPer], vw[kPer], vwn[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const size_t i = b0 + (size_t)k * kRedThreads + t;
    const size_t ic = (i < (size_t)W) ? i : 0;
    vwe[k] = dmc[2 * (size_t)W + ic];
    vw[k] = dmc[3 * (size_t)W + ic];
    vwn[k] = dmc[ic];
  }
  double s_we = 0.0, s_w = 0.0, s_wn = 0.0, mx = 0.0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (b0 + (size_t)k * kRedThreads + t < (size_t)W) {
      s_we += vwe[k];
      s_w += vw[k];
      s_wn += vwn[k];
      mx = fmax(mx, vwn[k]);
    }
  }
  sh[0][t] = s_we;
  sh[1][t] = s_w;
  sh[2][t] = s_wn;
  sh[3][t] = mx;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
      sh[0][t] += sh[0][t + off];
      sh[1][t] += sh[1][t + off];
      sh[2][t] += sh[2][t + off];
      sh[3][t] = fmax(sh[3][t], sh[3][t + off]);
    }
    __syncthreads();
  }
  if (t < 4) part[(size_t)t * n_parts + blockIdx.x] = sh[t][0];
}

__device__ void step_end(const DmcArgs &a);

// The second stage over the n_parts partials, in one fixed order (per-thread
// strided sums, then an LDS tree), leaving the four results in sh[q][0]; with
// `write` the step's ensemble energy (dmc.rs:117-121) and SRBrancher's mean
// weight and normalisation (branching.rs:21-24) go to the state, and the same
// thread then runs the step's bookkeeping (no later launch of the step reads
// what it changes: the branching reads its stream key from `branch`).
__device__ void finish_reduction(const DmcArgs &a, int n_parts, double (*sh)[kRedThreads], bool write) {
  const int t = threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = t; i < n_parts; i += kRedThreads) {
    v[0] += a.part[i];
    v[1] += a.part[(size_t)n_parts + i];
    v[2] += a.part[2 * (size_t)n_parts + i];
    v[3] = fmax(v[3], a.part[3 * (size_t)n_parts + i]);
  }
  for (int q = 0; q < 4; ++q) sh[q][t] = v[q];
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (t < off) {
      sh[0][t] += sh[0][t + off];
      sh[1][t] += sh[1][t + off];
      sh[2][t] += sh[2][t + off];
      sh[3][t] = fmax(sh[3][t], sh[3][t + off]);
    }
    __syncthreads();
  }
  if (write && t == 0) {
    DmcState *st = a.st;
    st->sum_we = sh[0][0];
    st->sum_w = sh[1][0];
    st->sum_wnew = sh[2][0];
    st->max_w = sh[3][0];
    const double ens = sh[0][0] / sh[1][0];
    if (st->it < a.cap_steps) a.step_energy[st->it] = ens;
    a.block_buf[st->j] = ens;
    st->gw = sh[2][0] / (double)a.W;
    st->norm = (double)a.W / sh[3][0];
    step_end(a);
  }
}

__global__ void __launch_bounds__(kRedThreads) reduce2_kernel(DmcArgs a, int n_parts) {
  __shared__ double sh[4][kRedThreads];
  finish_reduction(a, n_parts, sh, true);
}

// ---- inclusive scan of u64 counts (n from the host bound or a device count) ----

__global__ void __launch_bounds__(kScanThreads) scan_local_kernel(uint64_t *x, size_t n_max, const uint32_t *n_dev,
                                                                  uint64_t *bsum) {
  __shared__ uint64_t sh[kScanThreads];
  const size_t n = n_dev ? (size_t)*n_dev : n_max;
  const int t = threadIdx.x;
  const size_t e0 = (size_t)blockIdx.x * kScanBlock + (size_t)t * kScanPerThread;
  uint64_t v[kScanPerThread];
  uint64_t run = 0;
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = (e0 + k < n) ? x[e0 + k] : 0;
    run += v[k];
    v[k] = run;
  }
  sh[t] = run;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint64_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const uint64_t before = (t > 0) ? sh[t - 1] : 0;
  for (int k = 0; k < kScanPerThread; ++k)
    if (e0 + k < n) x[e0 + k] = v[k] + before;
  if (t == kScanThreads - 1) bsum[blockIdx.x] = sh[t];
}

// exclusive scan of the block sums, in place (one workgroup)
__global__ void __launch_bounds__(kScanThreads) scan_blocks_kernel(uint64_t *bsum, int nb) {
  __shared__ uint64_t sh[kScanThreads];
  const int t = threadIdx.x;
  const int per = (nb + kScanThreads - 1) / kScanThreads;
  const int b0 = t * per;
  uint64_t s = 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < nb) s += bsum[b0 + k];
  sh[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint64_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint64_t run = (t > 0) ? sh[t - 1] : 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < nb) {
      const uint64_t v = bsum[b0 + k];
      bsum[b0 + k] = run;
      run += v;
    }
}

__global__ void scan_add_kernel(uint64_t *x, size_t n_max, const uint32_t *n_dev, const uint64_t *bsum) {
  const size_t n = n_dev ? (size_t)*n_dev : n_max;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += bsum[i / kScanBlock];
}

hipError_t enqueue_scan(uint64_t *x, size_t n_max, const uint32_t *n_dev, uint64_t *bsum, hipStream_t stream) {
  const int nb = (int)dmc_scan_blocks(n_max);
  hipLaunchKernelGGL(scan_local_kernel, dim3(nb), dim3(kScanThreads), 0, stream, x, n_max, n_dev, bsum);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(kScanThreads), 0, stream, bsum, nb);
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n_max + 255) / 256)), dim3(256), 0, stream, x, n_max, n_dev,
                     (const uint64_t *)bsum);
  return hipGetLastError();
}

// ---- SRBrancher (branching.rs:15-44) ----

// Block-local inclusive scan of the integer weights (w W / max w) as u32,
// computed on the fly; block totals to bsum_a.  Every workgroup first
// finishes the ensemble reduction itself (the same partials in the same
// order, so the same W / max w everywhere); workgroup 0 stores the results.
__global__ void __launch_bounds__(kScanThreads) sr_scan_local_kernel(DmcArgs a, int n_parts) {
  static_assert(kScanThreads == kRedThreads, "one workgroup shape for both stages");
  __shared__ double red[4][kRedThreads];
  __shared__ uint64_t sh[kScanThreads];
  finish_reduction(a, n_parts, red, blockIdx.x == 0);
  const size_t n = (size_t)a.W;
  const double norm = (double)a.W / red[3][0];
  const int t = threadIdx.x;
  const size_t e0 = (size_t)blockIdx.x * kScanBlock + (size_t)t * kScanPerThread;
  uint64_t v[kScanPerThread];
  uint64_t run = 0;
  for (int k = 0; k < kScanPerThread; ++k) {
    run += (e0 + k < n) ? as_u32(a.dmc[e0 + k] * norm) : 0;
    v[k] = run;
  }
  sh[t] = run;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint64_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const uint64_t before = (t > 0) ? sh[t - 1] : 0;
  for (int k = 0; k < kScanPerThread; ++k)
    if (e0 + k < n) a.cnt_a[e0 + k] = v[k] + before;
  if (t == kScanThreads - 1) a.bsum_a[blockIdx.x] = sh[t];
}

// Walker s's positions, inverses and cached local energy into column j of
// the gathered copies, which the next step starts from
__device__ __forceinline__ void gather_walker(const DmcArgs &a, uint32_t j, uint32_t s) {
  const size_t W = (size_t)a.W;
  for (int r = 0; r < a.pos_rows; ++r) a.pos_s[r * W + j] = a.pos[r * W + s];
  if (a.b_elem == 4) {
    for (int r = 0; r < a.b_rows; ++r) ((float *)a.B_s)[r * W + j] = ((const float *)a.B)[r * W + s];
  } else {
    for (int r = 0; r < a.b_rows; ++r) ((double *)a.B_s)[r * W + j] = ((const double *)a.B)[r * W + s];
  }
  a.el_s[j] = a.dmc[W + s];
}

// WeightedChoice: the first index whose cumulative weight exceeds u.  The
// cumulative weight through walker k is its block-local sum plus the
// exclusive offset of its scan block.  With LOCAL_OFFSETS (at most
// kScanThreads scan blocks) every workgroup scans the block totals in LDS
// itself; otherwise bsum_a already holds the offsets.  Thread j then gathers
// its chosen walker into column j of the copies the next step starts from.
template <bool LOCAL_OFFSETS>
__global__ void __launch_bounds__(kScanThreads) sr_select_kernel(DmcArgs a) {
  __shared__ uint64_t off[kScanThreads];
  const int t = threadIdx.x;
  if constexpr (LOCAL_OFFSETS) {
    const int nb = (a.W + kScanBlock - 1) / kScanBlock;
    const uint64_t v = t < nb ? a.bsum_a[t] : 0;
    off[t] = v;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
      const uint64_t add = (t >= o) ? off[t - o] : 0;
      __syncthreads();
      off[t] += add;
      __syncthreads();
    }
    off[t] -= v;  // each thread its own entry: inclusive to exclusive
    __syncthreads();
  }
  auto cum = [&](uint32_t k) -> uint64_t {
    if constexpr (LOCAL_OFFSETS) return a.cnt_a[k] + off[k / kScanBlock];
    else return a.cnt_a[k] + a.bsum_a[k / kScanBlock];
  };
  const int j = blockIdx.x * blockDim.x + t;
  if (j >= a.W) return;
  DmcState *st = a.st;
  const uint64_t total = cum((uint32_t)a.W - 1);
  if (j == 0) st->total = total;
  if (total == 0) {  // WeightedChoice::new panics: flag it, keep the walkers
    if (j == 0) st->error = 1;
    a.src[j] = (uint32_t)j;
    a.wnew[j] = a.dmc[j];
    gather_walker(a, (uint32_t)j, (uint32_t)j);
    return;
  }
  Stream rs;
  rs.init(a.seed, a.sys, st->branch, kDmcBranchEpoch);
  const uint64_t u = uniform_below(rs.draw((uint64_t)j), total);
  uint32_t lo = 0, hi = (uint32_t)a.W - 1;
  if constexpr (LOCAL_OFFSETS) {
    // the scan block holding the answer first, from the offsets in LDS: the
    // first block b whose inclusive end off[b + 1] exceeds u (the last block
    // ends at the total, which exceeds u); before it the cumulative weight is
    // off[b] <= u, so the first index past u lies in block b and the global
    // search below covers one block's kScanBlock entries instead of all W
    const int nb = (a.W + kScanBlock - 1) / kScanBlock;
    int bl = 0, bh = nb - 1;
    while (bl < bh) {
      const int bm = bl + (bh - bl) / 2;
      if (off[bm + 1] > u) bh = bm;
      else bl = bm + 1;
    }
    lo = (uint32_t)bl * kScanBlock;
    hi = (uint32_t)min(a.W - 1, (bl + 1) * kScanBlock - 1);
  }
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (cum(mid) > u) hi = mid;
    else lo = mid + 1;
  }
  a.src[j] = lo;
  a.wnew[j] = st->gw;
  gather_walker(a, (uint32_t)j, lo);
}

// ---- SimpleBranching (branching.rs:57-91) ----

__global__ void simple_counts_kernel(DmcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.W) return;
  Stream rs;
  rs.init(a.seed, a.sys, a.st->branch, kDmcBranchEpoch);
  uint64_t n = as_usize(a.dmc[i] + u64_standard(rs.draw((uint64_t)i)));
  if (n > 3) n = 3;
  a.cnt_a[i] = n >= 1 ? 1 : 0;
  a.cnt_b[i] = n >= 1 ? n - 1 : 0;
}

// survivors in order, then the births in order
__global__ void simple_list_kernel(DmcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.W) return;
  const uint64_t S = a.cnt_a[a.W - 1], B = a.cnt_b[a.W - 1];
  if (i == 0) {
    a.st->S = (uint32_t)S;
    a.st->M = (uint32_t)(S + B);
  }
  const uint64_t ka = a.cnt_a[i], kb = a.cnt_b[i];
  const uint64_t keep = ka - (i ? a.cnt_a[i - 1] : 0);
  const uint64_t births = kb - (i ? a.cnt_b[i - 1] : 0);
  if (keep) a.list[ka - 1] = (uint32_t)i;
  for (uint64_t r = 0; r < births; ++r) a.list[S + kb - births + r] = (uint32_t)i;
}

// fewer than W: clones of uniformly chosen original walkers
__global__ void simple_clone_kernel(DmcArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t M = a.st->M;
  if (M >= (uint32_t)a.W || k >= a.W - (int)M) return;
  Stream rs;
  rs.init(a.seed, a.sys, a.st->branch, kDmcBranchEpoch);
  a.list[M + k] = (uint32_t)uniform_below(rs.draw((uint64_t)a.W + k), (uint64_t)a.W);
}

// more than W: M - W removals, each at a uniform position of the shrinking
// list.  The k-th removal's original position is the least p with
// p = r_k + #{earlier removals <= p} (iterated from p = r_k; p never passes
// M - 1, so the loop ends).
constexpr int kRemoveThreads = 1024;
__global__ void __launch_bounds__(kRemoveThreads) simple_remove_kernel(DmcArgs a) {
  __shared__ uint32_t sh[kRemoveThreads];
  const uint32_t M = a.st->M, W = (uint32_t)a.W;
  if (M <= W) return;
  const uint32_t E = M - W;
  const int t = threadIdx.x;
  Stream rs;
  rs.init(a.seed, a.sys, a.st->branch, kDmcBranchEpoch);
  for (uint32_t k = 0; k < E; ++k) {
    const uint64_t r = uniform_below(rs.draw((uint64_t)W + k), (uint64_t)(M - k));
    uint64_t p = r;
    for (;;) {
      uint32_t c = 0;
      for (uint32_t q = t; q < k; q += kRemoveThreads) c += (a.removed[q] <= p) ? 1u : 0u;
      sh[t] = c;
      __syncthreads();
      for (int off = kRemoveThreads / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
      }
      const uint64_t np = r + sh[0];
      __syncthreads();
      if (np == p) break;
      p = np;
    }
    if (t == 0) a.removed[k] = (uint32_t)p;
    __syncthreads();
  }
}

// flags of the list entries that survive the removals (cnt_a, M entries)
__global__ void simple_flags_kernel(DmcArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.st->M) a.cnt_a[i] = 1;
}
__global__ void simple_unflag_kernel(DmcArgs a) {
  const uint32_t M = a.st->M;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (M > (uint32_t)a.W && k < M - (uint32_t)a.W) a.cnt_a[a.removed[k]] = 0;
}

// the final W walkers; SimpleBranching keeps each walker's weight
__global__ void simple_final_kernel(DmcArgs a) {
  const uint32_t M = a.st->M, W = (uint32_t)a.W;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (M > W) {
    if (i < M && (a.cnt_a[i] - (i ? a.cnt_a[i - 1] : 0)) == 1) {
      const uint64_t j = a.cnt_a[i] - 1;
      const uint32_t s = a.list[i];
      a.src[j] = s;
      a.wnew[j] = a.dmc[s];
    }
  } else if (i < W) {
    const uint32_t s = a.list[i];
    a.src[i] = s;
    a.wnew[i] = a.dmc[s];
  }
}

// ---- walker state after branching ----

// SimpleBranching: positions, inverses and cached local energies of the
// chosen walkers into the gathered copies, one flat index over
// [3N positions | inverses | energy] x W; the next step starts from them.
__global__ void gather_all_kernel(DmcArgs a) {
  const size_t W = (size_t)a.W;
  const size_t np = (size_t)a.pos_rows * W, nb = (size_t)a.b_rows * W, n = np + nb + W;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    if (e < np) {
      const size_t r = e / W, j = e % W;
      a.pos_s[e] = a.pos[r * W + a.src[j]];
    } else if (e < np + nb) {
      const size_t f = e - np, r = f / W, j = f % W;
      if (a.b_elem == 4)
        ((float *)a.B_s)[f] = ((const float *)a.B)[r * W + a.src[j]];
      else
        ((double *)a.B_s)[f] = ((const double *)a.B)[r * W + a.src[j]];
    } else {
      const size_t j = e - np - nb;
      a.el_s[j] = a.dmc[W + a.src[j]];
    }
  }
}

// Between calls the walkers live in the walker state (positions, inverses,
// dmc rows 0-1: weight, cached local energy); within a call, between steps,
// in the gathered copies (weights in wnew).  to_main: the copies into the
// walker state (a call's end), else the walker state into the copies (its
// start).
__global__ void sync_state_kernel(DmcArgs a, int to_main) {
  const size_t W = (size_t)a.W;
  const size_t np = (size_t)a.pos_rows * W, nb = (size_t)a.b_rows * W, n = np + nb + 2 * W;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    if (e < np) {
      if (to_main) a.pos[e] = a.pos_s[e];
      else a.pos_s[e] = a.pos[e];
    } else if (e < np + nb) {
      const size_t f = e - np;
      if (a.b_elem == 4) {
        if (to_main) ((float *)a.B)[f] = ((const float *)a.B_s)[f];
        else ((float *)a.B_s)[f] = ((const float *)a.B)[f];
      } else {
        if (to_main) ((double *)a.B)[f] = ((const double *)a.B_s)[f];
        else ((double *)a.B_s)[f] = ((const double *)a.B)[f];
      }
    } else if (e < np + nb + W) {
      const size_t j = e - np - nb;
      if (to_main) a.dmc[W + j] = a.el_s[j];
      else a.el_s[j] = a.dmc[W + j];
    } else {
      const size_t j = e - np - nb - W;
      if (to_main) a.dmc[j] = a.wnew[j];
      else a.wnew[j] = a.dmc[j];
    }
  }
}

// update_energies (dmc.rs:155-202) at the end of a block; counters advance,
// the sweep counters too for the one system that carries them
__device__ void step_end(const DmcArgs &a) {
  if (a.sweeps_done) {
    *a.sweeps_done += 1;
    *a.trace_sweeps += 1;
  }
  DmcState *st = a.st;
  st->branch = (uint32_t)st->steps;
  st->steps += 1;
  st->it += 1;
  if (++st->j < a.block_size) return;
  st->j = 0;
  const uint32_t b = st->block_nr++;
  if (b < a.num_eq_blocks) return;
  double energy = 0.0;
  for (uint32_t k = 0; k < a.block_size; ++k) energy += a.block_buf[k];
  energy /= (double)a.block_size;
  const uint32_t o = st->n_out;
  if (b == a.num_eq_blocks) {
    st->eref = (st->eref + energy) / 2.0;
    a.energies[o] = energy;
    a.vars[o] = 0.0;
  } else {
    const double prev = a.energies[o - 1];
    const double k = (double)(b - a.num_eq_blocks);
    a.energies[o] = prev + (energy - prev) / k;
    st->eref = (st->eref + a.energies[o]) / 2.0;
    a.vars[o] = a.vars[o - 1] + ((energy - prev) * (energy - a.energies[o]) - a.vars[o - 1]) / k;
  }
  st->n_out = o + 1;
}

__global__ void state_set_kernel(DmcState *st, DmcState v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *st = v;
}

// a new DmcRunner::diffuse call: fresh block bookkeeping, E_ref and the
// branching stream carry on
__global__ void call_start_kernel(DmcState *st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->it = 0;
  st->j = 0;
  st->block_nr = 0;
  st->n_out = 0;
  st->error = 0;
}

unsigned grid_for(size_t n, unsigned cap = 65535u) {
  const size_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

size_t dmc_parts(int W) { return ((size_t)W + kRedChunk - 1) / kRedChunk; }
size_t dmc_scan_blocks(size_t n) { return (n + kScanBlock - 1) / kScanBlock; }

hipError_t launch_dmc_init_configs(double *pos, int N, int W, uint64_t seed, uint32_t sys, hipStream_t stream) {
  hipLaunchKernelGGL(init_configs_kernel, dim3((W + 255) / 256), dim3(256), 0, stream, pos, N, W, seed, sys);
  return hipGetLastError();
}

hipError_t launch_dmc_state_set(DmcState *st, const DmcState &v, hipStream_t stream) {
  hipLaunchKernelGGL(state_set_kernel, dim3(1), dim3(64), 0, stream, st, v);
  return hipGetLastError();
}

hipError_t launch_dmc_call_start(DmcState *st, hipStream_t stream) {
  hipLaunchKernelGGL(call_start_kernel, dim3(1), dim3(64), 0, stream, st);
  return hipGetLastError();
}

hipError_t launch_dmc_after_step(const DmcArgs &a, hipStream_t stream) {
  const int W = a.W;
  const int n_parts = (int)dmc_parts(W);
  const dim3 gw((W + 255) / 256), b256(256);
  hipLaunchKernelGGL(reduce1_kernel, dim3(n_parts), dim3(kRedThreads), 0, stream, (const double *)a.dmc, W, a.part,
                     n_parts);
  hipError_t e;
  if (a.brancher == kDmcBranchSR) {
    const int nb = (int)dmc_scan_blocks((size_t)W);
    hipLaunchKernelGGL(sr_scan_local_kernel, dim3(nb), dim3(kScanThreads), 0, stream, a, n_parts);
    if (nb <= kScanThreads) {
      hipLaunchKernelGGL(sr_select_kernel<true>, gw, dim3(kScanThreads), 0, stream, a);
    } else {
      hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(kScanThreads), 0, stream, a.bsum_a, nb);
      hipLaunchKernelGGL(sr_select_kernel<false>, gw, dim3(kScanThreads), 0, stream, a);
    }
  } else {
    hipLaunchKernelGGL(reduce2_kernel, dim3(1), dim3(kRedThreads), 0, stream, a, n_parts);
    hipLaunchKernelGGL(simple_counts_kernel, gw, b256, 0, stream, a);
    if ((e = enqueue_scan(a.cnt_a, (size_t)W, nullptr, a.bsum_a, stream)) != hipSuccess) return e;
    if ((e = enqueue_scan(a.cnt_b, (size_t)W, nullptr, a.bsum_b, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(simple_list_kernel, gw, b256, 0, stream, a);
    hipLaunchKernelGGL(simple_clone_kernel, gw, b256, 0, stream, a);
    hipLaunchKernelGGL(simple_remove_kernel, dim3(1), dim3(kRemoveThreads), 0, stream, a);
    const dim3 g3((unsigned)((3 * (size_t)W + 255) / 256));
    hipLaunchKernelGGL(simple_flags_kernel, g3, b256, 0, stream, a);
    hipLaunchKernelGGL(simple_unflag_kernel, dim3((unsigned)((2 * (size_t)W + 255) / 256)), b256, 0, stream, a);
    if ((e = enqueue_scan(a.cnt_a, 3 * (size_t)W, &a.st->M, a.bsum_a, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(simple_final_kernel, g3, b256, 0, stream, a);
    const size_t n_gather = (size_t)(a.pos_rows + a.b_rows + 1) * W;
    hipLaunchKernelGGL(gather_all_kernel, dim3(grid_for(n_gather)), b256, 0, stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_dmc_sync_state(const DmcArgs &a, bool to_main, hipStream_t stream) {
  const size_t n_all = (size_t)(a.pos_rows + a.b_rows + 2) * (size_t)a.W;
  hipLaunchKernelGGL(sync_state_kernel, dim3(grid_for(n_all)), dim3(256), 0, stream, a, to_main ? 1 : 0);
  return hipGetLastError();
}

}  // namespace mole
