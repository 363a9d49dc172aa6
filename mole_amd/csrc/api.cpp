// Host side of the C ABI declared in include/mole_gpu.h.
//
// A context holds, per system, the walker positions as [3N][W] doubles,
// the per-walker accumulators as [column][W] and the optional trace and
// sample buffers, plus the state every walker shares: the Philox epoch,
// the next draw index and the fill of the block in progress.  Arguments
// are validated before any HIP call, so invalid calls fail the same way
// with or without a device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mole_gpu.h"
#include "internal.hpp"

using namespace mole;

namespace {

thread_local std::string g_last_error;

mole_status fail(mole_status s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

mole_status ok() {
  g_last_error.clear();
  return MOLE_OK;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE, \
                  "%s: %s", #expr, hipGetErrorString(e_));                         \
  } while (0)

struct SysState {
  Sys2 sp{};
  uint32_t kind = 0, n_up = 0, n_down = 0, N = 0, n_params = 0;
  int W = 0;
  double *d_pos = nullptr;
  double *d_acc = nullptr;
  uint64_t *d_trace_mask = nullptr;
  double *d_trace_e = nullptr;
  double *d_samp_psi = nullptr;
  double *d_samp_g = nullptr;
  uint64_t n_samples = 0, n_blocks = 0, n_moves = 0;
  // column sums of d_acc, valid until the next launch that writes d_acc
  double cols[kAccColumns2] = {};
  bool cols_valid = false;
};

}  // namespace

struct mole_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mole_run_desc run{};
  std::vector<SysState> sys;
  uint32_t epoch = 0;
  uint64_t position = 0;
  uint32_t block_fill = 0;
  uint32_t trace_sweeps = 0, trace_samples = 0;
  double *d_red = nullptr;
  uint32_t last_launches = 0;
  bool timed = false;
  int variant = 0;        // MOLE_VARIANT_*
  int block_threads = 64; // one-lane-per-walker kernel
  uint32_t max_sweeps_per_launch = 4096;
};

namespace {

mole_status validate(const mole_system_desc *sys, uint32_t n_sys, const mole_run_desc *run) {
  if (!sys || !run) return fail(MOLE_ERR_INVALID_ARGUMENT, "null descriptor");
  if (n_sys == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "n_sys must be >= 1");
  if (run->proposal != MOLE_PROPOSAL_BOX)
    return fail(MOLE_ERR_UNSUPPORTED, "proposal %u not supported", run->proposal);
  // rand's Range::new asserts low < high (metrop.rs:66)
  if (!(run->step > 0.0) || !std::isfinite(run->step))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "box side must be a positive finite number");
  if (run->block_size == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "block_size must be >= 1");
  if (run->o_convention > MOLE_O_REFERENCE)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown O convention %u", run->o_convention);
  if ((run->flags & (MOLE_FLAG_TRACE | MOLE_FLAG_SAMPLES)) && run->record_sweeps == 0)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "record_sweeps must be > 0 with trace/sample flags");
  for (uint32_t s = 0; s < n_sys; ++s) {
    const mole_system_desc &d = sys[s];
    if (d.wf_kind != MOLE_WF_HELIUM_PADE && d.wf_kind != MOLE_WF_LCAO2_PADE)
      return fail(MOLE_ERR_UNSUPPORTED, "system %u: wavefunction kind %u not supported", s, d.wf_kind);
    if (d.hamiltonian > MOLE_H_IONIC)
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: unknown hamiltonian %u", s, d.hamiltonian);
    if (d.n_up != 1 || d.n_down != 1)
      return fail(MOLE_ERR_SHAPE, "system %u: two-electron kinds need n_up = n_down = 1", s);
    if (d.n_nuclei < 1 || d.n_nuclei > (uint32_t)kMaxNuc2)
      return fail(MOLE_ERR_SHAPE, "system %u: need 1..%d nuclei", s, kMaxNuc2);
    if (!d.nuclei_xyz || !d.nuclei_charge)
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: null nuclei", s);
    if (d.n_params != (uint32_t)kNParam2 || !d.params)
      return fail(MOLE_ERR_SHAPE, "system %u: expected %d parameters (zeta, b, c)", s, kNParam2);
    if (d.n_walkers == 0 || d.n_walkers > (1u << 26))
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: n_walkers must be in 1..2^26", s);
  }
  return MOLE_OK;
}

void fill_sys2(SysState &st, const mole_system_desc &d) {
  st.kind = d.wf_kind;
  st.n_up = d.n_up;
  st.n_down = d.n_down;
  st.N = d.n_up + d.n_down;
  st.n_params = d.n_params;
  st.W = (int)d.n_walkers;
  Sys2 &sp = st.sp;
  sp.kind = (int)d.wf_kind;
  sp.hamiltonian = (int)d.hamiltonian;
  sp.n_nuc = (int)d.n_nuclei;
  for (uint32_t I = 0; I < d.n_nuclei; ++I) {
    for (int k = 0; k < 3; ++k) sp.nuc[I][k] = d.nuclei_xyz[3 * I + k];
    sp.charge[I] = (double)d.nuclei_charge[I];
  }
  // nuclear repulsion as IonicPotential::new sums it (operator.rs:40-55)
  double rep = 0.0;
  for (uint32_t i = 0; i < d.n_nuclei; ++i)
    for (uint32_t j = i + 1; j < d.n_nuclei; ++j) {
      double dx = sp.nuc[j][0] - sp.nuc[i][0], dy = sp.nuc[j][1] - sp.nuc[i][1],
             dz = sp.nuc[j][2] - sp.nuc[i][2];
      double s2 = dx * dx;
      s2 = s2 + dy * dy;
      s2 = s2 + dz * dz;
      rep += (double)(d.nuclei_charge[i] * d.nuclei_charge[j]) / std::sqrt(s2);
    }
  sp.v_nn = rep;
  sp.zeta = d.params[0];
  sp.b = d.params[1];
  sp.c = d.params[2];
}

mole_status check_ctx(mole_ctx *ctx, uint32_t sys) {
  if (!ctx) return fail(MOLE_ERR_INVALID_ARGUMENT, "null context");
  if (sys >= ctx->sys.size()) return fail(MOLE_ERR_INVALID_ARGUMENT, "system index %u out of range", sys);
  return MOLE_OK;
}

void free_ctx(mole_ctx *ctx) {
  if (!ctx) return;
  for (SysState &st : ctx->sys) {
    if (st.d_pos) (void)hipFree(st.d_pos);
    if (st.d_acc) (void)hipFree(st.d_acc);
    if (st.d_trace_mask) (void)hipFree(st.d_trace_mask);
    if (st.d_trace_e) (void)hipFree(st.d_trace_e);
    if (st.d_samp_psi) (void)hipFree(st.d_samp_psi);
    if (st.d_samp_g) (void)hipFree(st.d_samp_g);
  }
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// Runs n sweeps; sweeps s >= sample_start of this call are sampled.
mole_status do_sweeps(mole_ctx *ctx, uint32_t n, uint32_t sample_start) {
  if (n == 0) return ok();
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  const double lo = -0.5 * ctx->run.step, hi = 0.5 * ctx->run.step;
  const double scale = hi - lo;
  const double offset = lo - scale;
  uint32_t launches = 0;
  uint32_t done = 0;
  while (done < n) {
    const uint32_t chunk = std::min(n - done, ctx->max_sweeps_per_launch);
    const uint32_t ss = (sample_start > done) ? std::min(sample_start - done, chunk) : 0;
    const uint32_t n_samp = chunk - ss;
    for (uint32_t s = 0; s < ctx->sys.size(); ++s) {
      SysState &st = ctx->sys[s];
      Sweep2Args a{};
      a.sp = st.sp;
      a.W = st.W;
      a.pos = st.d_pos;
      a.acc = st.d_acc;
      a.seed = ctx->run.seed;
      a.sys = s;
      a.epoch = ctx->epoch;
      a.draw0 = ctx->position;
      a.scale = scale;
      a.offset = offset;
      a.n_sweeps = chunk;
      a.sample_start = ss;
      a.block_size = ctx->run.block_size;
      a.block_fill = ctx->block_fill;
      a.o_conv = (int)ctx->run.o_convention;
      a.trace_mask = st.d_trace_mask;
      a.trace_e = st.d_trace_e;
      a.samp_psi = st.d_samp_psi;
      a.samp_g = st.d_samp_g;
      a.trace_sweep0 = ctx->trace_sweeps;
      a.trace_sample0 = ctx->trace_samples;
      a.cap_sweeps = (st.d_trace_mask) ? ctx->run.record_sweeps : 0;
      a.cap_samples = (st.d_trace_e || st.d_samp_psi) ? ctx->run.record_sweeps : 0;
      st.cols_valid = false;
      HIP_TRY(launch_sweep2(a, ctx->variant, ctx->block_threads, ctx->stream));
      ++launches;
      const uint64_t filled = (uint64_t)ctx->block_fill + n_samp;
      st.n_samples += (uint64_t)st.W * n_samp;
      st.n_blocks += (uint64_t)st.W * (filled / ctx->run.block_size);
      st.n_moves += (uint64_t)st.W * st.N * chunk;
    }
    ctx->position += (uint64_t)chunk * ctx->sys[0].N * 4;
    ctx->block_fill = (uint32_t)(((uint64_t)ctx->block_fill + n_samp) % ctx->run.block_size);
    ctx->trace_sweeps += chunk;
    ctx->trace_samples += n_samp;
    done += chunk;
  }
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  ctx->last_launches = launches;
  ctx->timed = true;
  return ok();
}

}  // namespace

extern "C" {

const char *mole_status_string(mole_status s) {
  switch (s) {
    case MOLE_OK: return "ok";
    case MOLE_ERR_LINALG: return "linear algebra error";
    case MOLE_ERR_SHAPE: return "shape error";
    case MOLE_ERR_FUNC: return "function error";
    case MOLE_ERR_OPERATOR_VALUE_ACCESS: return "operator value access error";
    case MOLE_ERR_DATA_ACCESS: return "data access error";
    case MOLE_ERR_EMPTY_CACHE: return "empty cache error";
    case MOLE_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MOLE_ERR_DEVICE: return "device error";
    case MOLE_ERR_UNSUPPORTED: return "unsupported";
    case MOLE_ERR_OUT_OF_MEMORY: return "out of device memory";
    default: return "unknown status";
  }
}

const char *mole_last_error(void) { return g_last_error.c_str(); }

mole_status mole_device_info(char *name, size_t cap, int32_t *cu_count) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (name && cap) {
    std::snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return ok();
}

mole_status mole_ctx_create(const mole_system_desc *sys, uint32_t n_sys, const mole_run_desc *run,
                            mole_ctx **out) {
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  mole_status st = validate(sys, n_sys, run);
  if (st != MOLE_OK) return st;
  mole_ctx *ctx = new mole_ctx();
  ctx->run = *run;
  ctx->sys.resize(n_sys);
  for (uint32_t s = 0; s < n_sys; ++s) fill_sys2(ctx->sys[s], sys[s]);
  auto bail = [&](mole_status s) {
    std::string msg = g_last_error;
    free_ctx(ctx);
    g_last_error = msg;
    return s;
  };
#define CTX_TRY(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return bail(fail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE, \
                       "%s: %s", #expr, hipGetErrorString(e_)));                    \
  } while (0)
  CTX_TRY(hipGetDevice(&ctx->device));
  CTX_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  CTX_TRY(hipEventCreate(&ctx->ev0));
  CTX_TRY(hipEventCreate(&ctx->ev1));
  CTX_TRY(hipMalloc(&ctx->d_red, sizeof(double) * kAccColumns2));
  for (SysState &s : ctx->sys) {
    const size_t W = (size_t)s.W;
    CTX_TRY(hipMalloc(&s.d_pos, sizeof(double) * 3 * s.N * W));
    CTX_TRY(hipMalloc(&s.d_acc, sizeof(double) * kAccColumns2 * W));
    CTX_TRY(launch_fill(s.d_pos, 3 * s.N * W, 0.0, ctx->stream));
    CTX_TRY(launch_fill(s.d_acc, kAccColumns2 * W, 0.0, ctx->stream));
    const size_t cap = run->record_sweeps;
    if (run->flags & MOLE_FLAG_TRACE) {
      CTX_TRY(hipMalloc(&s.d_trace_mask, sizeof(uint64_t) * cap * W));
      CTX_TRY(hipMalloc(&s.d_trace_e, sizeof(double) * cap * W));
      CTX_TRY(hipMemsetAsync(s.d_trace_mask, 0, sizeof(uint64_t) * cap * W, ctx->stream));
      CTX_TRY(launch_fill(s.d_trace_e, cap * W, 0.0, ctx->stream));
    }
    if (run->flags & MOLE_FLAG_SAMPLES) {
      CTX_TRY(hipMalloc(&s.d_samp_psi, sizeof(double) * cap * W));
      CTX_TRY(hipMalloc(&s.d_samp_g, sizeof(double) * cap * W * kNParam2));
      if (!s.d_trace_e) {
        CTX_TRY(hipMalloc(&s.d_trace_e, sizeof(double) * cap * W));
        CTX_TRY(launch_fill(s.d_trace_e, cap * W, 0.0, ctx->stream));
      }
    }
  }
  CTX_TRY(hipStreamSynchronize(ctx->stream));
#undef CTX_TRY
  *out = ctx;
  return ok();
}

void mole_ctx_destroy(mole_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
}

mole_status mole_init_walkers(mole_ctx *ctx) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  for (uint32_t s = 0; s < ctx->sys.size(); ++s) {
    SysState &ss = ctx->sys[s];
    HIP_TRY(launch_init2(ss.d_pos, ss.W, ctx->run.seed, s, ctx->epoch, ctx->stream));
  }
  ctx->position = 3ull * ctx->sys[0].N;
  return ok();
}

mole_status mole_set_configs(mole_ctx *ctx, uint32_t sys, const double *cfg, size_t n_walkers) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null cfg");
  if (n_walkers != (size_t)s.W) return fail(MOLE_ERR_SHAPE, "expected %d walkers", s.W);
  const size_t D = 3 * s.N, W = s.W;
  std::vector<double> soa(D * W);
  for (size_t w = 0; w < W; ++w)
    for (size_t k = 0; k < D; ++k) soa[k * W + w] = cfg[w * D + k];
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(s.d_pos, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ok();
}

mole_status mole_get_configs(mole_ctx *ctx, uint32_t sys, double *cfg, size_t n_walkers) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null cfg");
  if (n_walkers != (size_t)s.W) return fail(MOLE_ERR_SHAPE, "expected %d walkers", s.W);
  const size_t D = 3 * s.N, W = s.W;
  std::vector<double> soa(D * W);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(soa.data(), s.d_pos, soa.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (size_t w = 0; w < W; ++w)
    for (size_t k = 0; k < D; ++k) cfg[w * D + k] = soa[k * W + w];
  return ok();
}

mole_status mole_broadcast_config(mole_ctx *ctx, uint32_t sys, const double *cfg) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null cfg");
  SysState &s = ctx->sys[sys];
  HIP_TRY(hipSetDevice(ctx->device));
  for (uint32_t k = 0; k < 3 * s.N; ++k)
    HIP_TRY(launch_fill(s.d_pos + (size_t)k * s.W, (size_t)s.W, cfg[k], ctx->stream));
  return ok();
}

mole_status mole_set_stream(mole_ctx *ctx, uint32_t epoch, uint64_t position) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  ctx->epoch = epoch;
  ctx->position = position;
  return ok();
}

mole_status mole_get_stream(mole_ctx *ctx, uint32_t *epoch, uint64_t *position) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (epoch) *epoch = ctx->epoch;
  if (position) *position = ctx->position;
  return ok();
}

mole_status mole_set_params(mole_ctx *ctx, uint32_t sys, const double *p, uint32_t n_params) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!p || n_params != s.n_params) return fail(MOLE_ERR_SHAPE, "expected %u parameters", s.n_params);
  s.sp.zeta = p[0];
  s.sp.b = p[1];
  s.sp.c = p[2];
  return ok();
}

mole_status mole_get_params(mole_ctx *ctx, uint32_t sys, double *p, uint32_t n_params) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!p || n_params != s.n_params) return fail(MOLE_ERR_SHAPE, "expected %u parameters", s.n_params);
  p[0] = s.sp.zeta;
  p[1] = s.sp.b;
  p[2] = s.sp.c;
  return ok();
}

mole_status mole_sweep(mole_ctx *ctx, uint32_t n_sweeps, uint32_t flags) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  return do_sweeps(ctx, n_sweeps, (flags & MOLE_SWEEP_SAMPLE) ? 0u : n_sweeps);
}

mole_status mole_run(mole_ctx *ctx, uint32_t steps, uint32_t block_size) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (block_size != ctx->run.block_size)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "block_size %u differs from the context's %u", block_size,
                ctx->run.block_size);
  if (steps < 2 * block_size)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "steps (%u) < 2 * block_size (%u)", steps, block_size);
  if (ctx->block_fill != 0) {
    // a run starts its own blocks: drop the partial block of earlier calls
    HIP_TRY(hipSetDevice(ctx->device));
    for (SysState &s : ctx->sys)
    {
      s.cols_valid = false;
      HIP_TRY(launch_fill(s.d_acc + (size_t)kAccBlockSum * s.W, (size_t)s.W, 0.0, ctx->stream));
    }
    ctx->block_fill = 0;
  }
  const uint32_t blocks = steps / block_size;
  return do_sweeps(ctx, blocks * block_size, block_size);
}

mole_status mole_reset_stats(mole_ctx *ctx) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  for (SysState &s : ctx->sys) {
    HIP_TRY(launch_fill(s.d_acc, (size_t)kAccColumns2 * s.W, 0.0, ctx->stream));
    s.n_samples = s.n_blocks = s.n_moves = 0;
    s.cols_valid = false;
  }
  ctx->block_fill = 0;
  ctx->trace_sweeps = ctx->trace_samples = 0;
  return ok();
}

static mole_status reduce_sys(mole_ctx *ctx, uint32_t sys, double *cols) {
  SysState &s = ctx->sys[sys];
  if (!s.cols_valid) {
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launch_reduce_columns(s.d_acc, s.W, kAccColumns2, ctx->d_red, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s.cols, ctx->d_red, sizeof(double) * kAccColumns2,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    s.cols_valid = true;
  }
  std::memcpy(cols, s.cols, sizeof s.cols);
  return MOLE_OK;
}

mole_status mole_read_stats(mole_ctx *ctx, uint32_t sys, mole_stats *out) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  double cols[kAccColumns2];
  st = reduce_sys(ctx, sys, cols);
  if (st != MOLE_OK) return st;
  const SysState &s = ctx->sys[sys];
  out->n_samples = s.n_samples;
  out->n_blocks = s.n_blocks;
  out->n_moves = s.n_moves;
  out->n_accepted = (uint64_t)cols[kAccAccepted];
  out->sum_e = cols[kAccSumE];
  out->sum_e2 = cols[kAccSumE2];
  out->sum_block_mean = cols[kAccSumBm];
  out->sum_block_mean2 = cols[kAccSumBm2];
  return ok();
}

mole_status mole_read_sr_sums(mole_ctx *ctx, uint32_t sys, mole_sr_sums *out) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  double cols[kAccColumns2];
  st = reduce_sys(ctx, sys, cols);
  if (st != MOLE_OK) return st;
  std::memset(out, 0, sizeof *out);
  const int np = kNParam2;
  out->n_params = np;
  out->n = ctx->sys[sys].n_samples;
  out->sum_e = cols[kAccSumE];
  int t = 0;
  for (int k = 0; k < np; ++k) {
    out->sum_o[k] = cols[kAccSumO + k];
    out->sum_oe[k] = cols[kAccSumOE + k];
    for (int l = k; l < np; ++l, ++t) {
      out->sum_oo[k * np + l] = cols[kAccSumOO + t];
      out->sum_oo[l * np + k] = cols[kAccSumOO + t];
    }
  }
  return ok();
}

mole_status mole_read_trace(mole_ctx *ctx, uint32_t sys, uint64_t *accept_mask, double *e_local,
                            size_t walkers, size_t sweeps, size_t samples) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!(ctx->run.flags & MOLE_FLAG_TRACE)) return fail(MOLE_ERR_DATA_ACCESS, "context has no trace");
  const size_t cap = ctx->run.record_sweeps;
  if (walkers != (size_t)s.W || sweeps > cap || samples > cap)
    return fail(MOLE_ERR_SHAPE, "trace request exceeds %d walkers x %zu records", s.W, cap);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (accept_mask) {
    std::vector<uint64_t> h(cap * s.W);
    HIP_TRY(hipMemcpy(h.data(), s.d_trace_mask, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(accept_mask + w * sweeps, h.data() + w * cap, sweeps * sizeof(uint64_t));
  }
  if (e_local) {
    std::vector<double> h(cap * s.W);
    HIP_TRY(hipMemcpy(h.data(), s.d_trace_e, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(e_local + w * samples, h.data() + w * cap, samples * sizeof(double));
  }
  return ok();
}

mole_status mole_read_samples(mole_ctx *ctx, uint32_t sys, double *e_local, double *psi,
                              double *pgrad, size_t walkers, size_t samples) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!(ctx->run.flags & MOLE_FLAG_SAMPLES)) return fail(MOLE_ERR_DATA_ACCESS, "context keeps no samples");
  const size_t cap = ctx->run.record_sweeps;
  if (walkers != (size_t)s.W || samples > cap)
    return fail(MOLE_ERR_SHAPE, "sample request exceeds %d walkers x %zu records", s.W, cap);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::vector<double> h(cap * s.W * kNParam2);
  if (e_local) {
    HIP_TRY(hipMemcpy(h.data(), s.d_trace_e, cap * s.W * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(e_local + w * samples, h.data() + w * cap, samples * sizeof(double));
  }
  if (psi) {
    HIP_TRY(hipMemcpy(h.data(), s.d_samp_psi, cap * s.W * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(psi + w * samples, h.data() + w * cap, samples * sizeof(double));
  }
  if (pgrad) {
    HIP_TRY(hipMemcpy(h.data(), s.d_samp_g, cap * s.W * kNParam2 * sizeof(double),
                      hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(pgrad + w * samples * kNParam2, h.data() + w * cap * kNParam2,
                  samples * kNParam2 * sizeof(double));
  }
  return ok();
}

mole_status mole_eval(mole_ctx *ctx, uint32_t sys, size_t n, const double *cfg, double *value,
                      double *grad, double *lap, double *e_local, double *pgrad) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null cfg");
  if (n == 0) return ok();
  SysState &s = ctx->sys[sys];
  const size_t D = 3 * s.N;
  HIP_TRY(hipSetDevice(ctx->device));
  double *d = nullptr;
  const size_t per = D + 1 + D + 1 + 1 + kNParam2;
  HIP_TRY(hipMalloc(&d, sizeof(double) * per * n));
  Eval2Args a{};
  a.sp = s.sp;
  a.n = n;
  double *d_cfg = d;
  a.cfg = d_cfg;
  a.value = value ? d + D * n : nullptr;
  a.grad = grad ? d + (D + 1) * n : nullptr;
  a.lap = lap ? d + (2 * D + 1) * n : nullptr;
  a.e_local = e_local ? d + (2 * D + 2) * n : nullptr;
  a.pgrad = pgrad ? d + (2 * D + 3) * n : nullptr;
  hipError_t e = hipMemcpyAsync(d_cfg, cfg, sizeof(double) * D * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_eval2(a, ctx->stream);
  auto back = [&](double *host, const double *dev, size_t cnt) {
    if (e == hipSuccess && host)
      e = hipMemcpyAsync(host, dev, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream);
  };
  back(value, a.value, n);
  back(grad, a.grad, D * n);
  back(lap, a.lap, n);
  back(e_local, a.e_local, n);
  back(pgrad, a.pgrad, kNParam2 * n);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_eval: %s", hipGetErrorString(e));
  return ok();
}

mole_status mole_compute_parameter_update(const mole_sr_sums *sums, uint32_t optimizer,
                                          double step, uint32_t sr_mode, double sr_eps,
                                          double *dp, double *g_out) {
  if (!sums || !dp) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  const int np = (int)sums->n_params;
  if (np < 1 || np > MOLE_MAX_PARAMS) return fail(MOLE_ERR_SHAPE, "n_params out of range");
  if (sums->n == 0) return fail(MOLE_ERR_DATA_ACCESS, "no samples");
  if (optimizer > MOLE_OPT_SR || sr_mode > MOLE_SR_REFERENCE)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown optimizer or SR mode");
  const double n = (double)sums->n;
  const double ebar = sums->sum_e / n;
  double g[MOLE_MAX_PARAMS], obar[MOLE_MAX_PARAMS];
  for (int i = 0; i < np; ++i) {
    obar[i] = sums->sum_o[i] / n;
    g[i] = 2.0 * (sums->sum_oe[i] - ebar * sums->sum_o[i]) / n;
  }
  if (g_out)
    for (int i = 0; i < np; ++i) g_out[i] = g[i];
  if (optimizer == MOLE_OPT_SD) {
    for (int i = 0; i < np; ++i) dp[i] = -(step * g[i]);
    return ok();
  }
  // augmented matrix [S | -g/2], Gaussian elimination with partial pivoting
  double A[MOLE_MAX_PARAMS][MOLE_MAX_PARAMS + 1];
  double osum = 0.0;
  for (int i = 0; i < np; ++i) osum += obar[i];
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      double s = sums->sum_oo[i * np + j] / n;
      s -= (sr_mode == MOLE_SR_REFERENCE) ? osum * osum : obar[i] * obar[j];
      A[i][j] = s;
    }
    A[i][i] *= 1.0 + ((sr_mode == MOLE_SR_REFERENCE) ? 1e-2 : sr_eps);
    A[i][np] = -0.5 * g[i];
  }
  for (int k = 0; k < np; ++k) {
    int piv = k;
    for (int i = k + 1; i < np; ++i)
      if (std::fabs(A[i][k]) > std::fabs(A[piv][k])) piv = i;
    if (!(std::fabs(A[piv][k]) > 0.0)) return fail(MOLE_ERR_LINALG, "SR matrix is singular");
    if (piv != k)
      for (int j = 0; j <= np; ++j) std::swap(A[k][j], A[piv][j]);
    for (int i = k + 1; i < np; ++i) {
      const double f = A[i][k] / A[k][k];
      for (int j = k; j <= np; ++j) A[i][j] -= f * A[k][j];
    }
  }
  for (int i = np - 1; i >= 0; --i) {
    double s = A[i][np];
    for (int j = i + 1; j < np; ++j) s -= A[i][j] * dp[j];
    dp[i] = s / A[i][i];
  }
  for (int i = 0; i < np; ++i) dp[i] = step * dp[i];
  return ok();
}

mole_status mole_set_launch_config(mole_ctx *ctx, uint32_t variant, uint32_t block_threads) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (variant > MOLE_VARIANT_WAVESPEC) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown variant %u", variant);
  if (block_threads != 0 && (block_threads % 64 != 0 || block_threads > 256))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "block_threads must be 0 or a multiple of 64 up to 256");
  if (variant == MOLE_VARIANT_WAVESPEC)
    for (const SysState &s : ctx->sys)
      if (!sweep2_wavespec_eligible(s.sp))
        return fail(MOLE_ERR_UNSUPPORTED, "wave-specialised kernel needs <= 2 nuclei (1 for helium)");
  ctx->variant = (int)variant;
  if (block_threads) ctx->block_threads = (int)block_threads;
  return ok();
}

mole_status mole_synchronize(mole_ctx *ctx) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ok();
}

mole_status mole_last_timing(mole_ctx *ctx, double *ms, uint32_t *launches) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (!ctx->timed) return fail(MOLE_ERR_EMPTY_CACHE, "no sweep has run");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  float f = 0.0f;
  HIP_TRY(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  if (ms) *ms = (double)f;
  if (launches) *launches = ctx->last_launches;
  return ok();
}

}  // extern "C"
