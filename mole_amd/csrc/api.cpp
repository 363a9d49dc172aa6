// Host side of the C ABI declared in include/mole_gpu.h.
//
// A context holds, per system, the walker positions as [3N][W] doubles,
// the per-walker accumulators as [column][W], optional trace and sample
// buffers and, for Slater-Jastrow systems, the inverse Slater matrices as
// [n_up^2 + n_dn^2][W] plus the system's basis/orbital tables.  The state
// every walker of a system shares -- Philox epoch, stream position, fill of
// the block in progress -- stays on the host and goes to the kernels as
// arguments.  Arguments are validated before any HIP call, so invalid calls
// fail the same way with or without a device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mole_gpu.h"
#include "internal.hpp"
#include "mlp_internal.hpp"
#include "sr_math.hpp"
#include "sj_internal.hpp"
#include "dmc_internal.hpp"

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

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE, "%s: %s", \
                  #expr, hipGetErrorString(e_));                                              \
  } while (0)

enum Family { kFamTwo = 0, kFamSJ = 1 };

struct SysState {
  uint32_t kind = 0, hamiltonian = 0, n_up = 0, n_down = 0, N = 0, n_params = 0;
  int W = 0;
  double *d_pos = nullptr;
  double *d_acc = nullptr;
  int n_cols = 0;
  uint64_t *d_trace_mask = nullptr;
  double *d_trace_e = nullptr;
  double *d_samp_psi = nullptr;
  double *d_samp_g = nullptr;
  uint64_t n_samples = 0, n_blocks = 0, n_moves = 0;
  std::vector<double> cols;  // column reductions of d_acc
  bool cols_valid = false;
  // two-electron kinds
  Sys2 sp{};
  // Slater-Jastrow
  std::vector<double> params;
  int n_nuc = 0, n_basis = 0, n_orb = 0, L = 0;
  size_t lds_walker = 0;
  double *d_tab = nullptr;  // nuclei xyz, charges, zeta, coefficients
  int *d_center = nullptr;
  void *d_B = nullptr;
  double v_nn = 0.0;
  uint64_t draw_base = 0, move_base = 0;
  bool any_gto = false, any_p = false;
  std::vector<double> zeta0;  // basis exponents as given (scaled by params[5] with six parameters)
  double *d_zeta0 = nullptr;  // the same on the device (the device optimisation loop rescales)
  double *d_ld = nullptr;  // [kSJLdColumns][W] log-determinant state
  // local operators: ops[0] the custom Hamiltonian, ops[1..] the observables
  std::vector<mole_operator> ops = std::vector<mole_operator>(1);
  bool custom_h = false;
  int n_obs = 0;
  mole_operator *d_ops = nullptr;
  double *d_obs_acc = nullptr;  // [2 MOLE_MAX_OBSERVABLES][W]
  // diffusion Monte Carlo (allocated by mole_dmc_init)
  struct Dmc {
    double *dmc = nullptr;  // [4][W]
    DmcState *st = nullptr;
    double *part = nullptr;
    uint64_t *cnt_a = nullptr, *cnt_b = nullptr, *bsum_a = nullptr, *bsum_b = nullptr;
    uint32_t *list = nullptr, *removed = nullptr, *src = nullptr;
    double *wnew = nullptr, *pos_s = nullptr, *el_s = nullptr;
    void *B_s = nullptr;
    double *step_energy = nullptr, *block_buf = nullptr, *energies = nullptr, *vars = nullptr;
    size_t cap_steps = 0, cap_block = 0, cap_blocks = 0;
    size_t steps_out = 0;  // steps recorded by the last mole_dmc_diffuse
  } dmc;
};

void free_dmc(SysState::Dmc &d) {
  void *ptrs[] = {d.dmc, d.st, d.part, d.cnt_a, d.cnt_b, d.bsum_a, d.bsum_b, d.list, d.removed, d.src,
                  d.wnew, d.pos_s, d.el_s, d.B_s, d.step_energy, d.block_buf, d.energies, d.vars};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  d = SysState::Dmc{};
}

}  // namespace

struct mole_ctx {
  int family = kFamTwo;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mole_run_desc run{};
  std::vector<SysState> sys;
  uint32_t epoch = 0;
  uint64_t position = 0;     // two-electron family: next draw index
  uint64_t sweeps_done = 0;  // Slater-Jastrow family: sweeps since the stream origin
  uint32_t block_fill = 0;
  uint32_t trace_sweeps = 0, trace_samples = 0;
  double *d_red = nullptr;
  uint32_t last_launches = 0;
  bool timed = false;
  int variant = 0;
  int block_threads = 64;
  uint32_t max_sweeps_per_launch = 4096;
  // Slater-Jastrow launch state
  SJSysDev *d_sysdev = nullptr;
  SJLaunchState *d_state = nullptr;
  uint32_t generation = 0;  // bumps when tiles/launch shape change (graphs go stale)
  SJTile *d_tiles = nullptr;
  int n_tiles = 0;
  size_t lds_bytes = 0, lds_limit = 65536;
  // tiles of one LDS footprint class (4 KiB steps) are launched together with
  // that class's dynamic LDS, so small walkers are not held to the largest
  // footprint's occupancy
  struct TileGroup {
    int first, count;
    size_t lds;
  };
  std::vector<TileGroup> tile_groups;
  // persistent schedule: every tile in one launch, longest first, taken from
  // a device ticket by `persist_wg` resident workgroups (0: CUs x occupancy)
  int schedule = MOLE_SCHEDULE_CLASSES;
  uint32_t persist_wg_req = 0;
  int persist_wg = 0;
  SJTile *d_tiles_lpt = nullptr;
  uint32_t *d_ticket = nullptr;
  int cu_count = 0;
  bool sysdev_dirty = true;
  mole_mlp *mlp = nullptr;  // network of the learned proposal (not owned)
  // device-side optimisation: parameters read by the kernels from d_params
  double *d_params = nullptr, *d_optlog = nullptr;
  mole_optimizer *d_opt = nullptr;
  size_t optlog_cap = 0;
  bool dev_params = false;
  // the next two-electron launch starts every walker's accumulators from zero
  // (a reset folded into the sweep kernel); cleared once a launch used it
  bool fresh_acc = false;
  // diffusion Monte Carlo
  bool dmc_ready = false;
  uint32_t dmc_brancher = 0;
};

// The learned proposal's network, its Adam state and its training data.
struct mole_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
  bf16x4 *d_fwd1 = nullptr, *d_bwd = nullptr;
  bf16x8 *d_fwd = nullptr;
  float *d_bias = nullptr;
  int adam_step = 0;
  // training configurations [n_cfg][3N] of one electron count
  double *d_cfg = nullptr;
  size_t n_cfg = 0, cap_cfg = 0;
  uint32_t data_N = 0;
  // per-step scratch
  float *d_partial = nullptr, *d_mid = nullptr;
  double *d_obj = nullptr, *d_obj_steps = nullptr, *d_mid_obj = nullptr;
  int partial_groups = 0, obj_steps_cap = 0;
  MlpDev net() const { return MlpDev{d_fwd1, d_fwd, d_bwd, d_bias}; }
};

namespace {

int family_of(uint32_t kind) { return kind == MOLE_WF_SLATER_JASTROW ? kFamSJ : kFamTwo; }

double nuclear_repulsion(const double *xyz, const int32_t *Z, uint32_t n) {
  // IonicPotential::new (operator.rs:40-55)
  double rep = 0.0;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j) {
      double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1], dz = xyz[3 * j + 2] - xyz[3 * i + 2];
      double s2 = dx * dx;
      s2 = s2 + dy * dy;
      s2 = s2 + dz * dz;
      rep += (double)(Z[i] * Z[j]) / std::sqrt(s2);
    }
  return rep;
}

mole_status validate(const mole_system_desc *sys, uint32_t n_sys, const mole_run_desc *run) {
  if (!sys || !run) return fail(MOLE_ERR_INVALID_ARGUMENT, "null descriptor");
  if (n_sys == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "n_sys must be >= 1");
  if (run->proposal != MOLE_PROPOSAL_BOX && run->proposal != MOLE_PROPOSAL_DIFFUSE &&
      run->proposal != MOLE_PROPOSAL_LEARNED)
    return fail(MOLE_ERR_UNSUPPORTED, "proposal %u not supported", run->proposal);
  // rand's Range::new asserts low < high (metrop.rs:66)
  if (run->proposal != MOLE_PROPOSAL_LEARNED && (!(run->step > 0.0) || !std::isfinite(run->step)))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "box side / time step must be a positive finite number");
  if (run->block_size == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "block_size must be >= 1");
  if (run->o_convention > MOLE_O_REFERENCE)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown O convention %u", run->o_convention);
  if ((run->flags & (MOLE_FLAG_TRACE | MOLE_FLAG_SAMPLES)) && run->record_sweeps == 0)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "record_sweeps must be > 0 with trace/sample flags");
  if (run->precision > MOLE_PREC_MIXED)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown precision %u", run->precision);
  const int fam = family_of(sys[0].wf_kind);
  for (uint32_t s = 0; s < n_sys; ++s) {
    const mole_system_desc &d = sys[s];
    if (d.wf_kind != MOLE_WF_HELIUM_PADE && d.wf_kind != MOLE_WF_LCAO2_PADE &&
        d.wf_kind != MOLE_WF_SLATER_JASTROW)
      return fail(MOLE_ERR_UNSUPPORTED, "system %u: wavefunction kind %u not supported", s, d.wf_kind);
    if (family_of(d.wf_kind) != fam)
      return fail(MOLE_ERR_UNSUPPORTED, "system %u: a context holds either two-electron kinds or "
                  "Slater-Jastrow systems", s);
    if (d.hamiltonian > MOLE_H_IONIC)
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: unknown hamiltonian %u", s, d.hamiltonian);
    if (!d.nuclei_xyz || !d.nuclei_charge || !d.params)
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: null nuclei or parameters", s);
    if (d.n_walkers == 0 || d.n_walkers > (1u << 26))
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: n_walkers must be in 1..2^26", s);
    if (fam == kFamTwo) {
      if (d.n_up != 1 || d.n_down != 1)
        return fail(MOLE_ERR_SHAPE, "system %u: two-electron kinds need n_up = n_down = 1", s);
      if (d.n_nuclei < 1 || d.n_nuclei > (uint32_t)kMaxNuc2)
        return fail(MOLE_ERR_SHAPE, "system %u: need 1..%d nuclei", s, kMaxNuc2);
      if (d.n_params != (uint32_t)kNParam2)
        return fail(MOLE_ERR_SHAPE, "system %u: expected %d parameters (zeta, b, c)", s, kNParam2);
    } else {
      if (d.n_up < 1 || d.n_up > (uint32_t)kSJMaxSpin || d.n_down > (uint32_t)kSJMaxSpin ||
          d.n_up + d.n_down > (uint32_t)kSJMaxElectrons)
        return fail(MOLE_ERR_SHAPE, "system %u: need 1 <= n_up <= %d, n_down <= %d, N <= %d", s,
                    kSJMaxSpin, kSJMaxSpin, kSJMaxElectrons);
      if (d.n_nuclei < 1 || d.n_nuclei > MOLE_MAX_NUCLEI)
        return fail(MOLE_ERR_SHAPE, "system %u: need 1..%d nuclei", s, MOLE_MAX_NUCLEI);
      if (d.n_params != 5 && d.n_params != (uint32_t)kSJNParam)
        return fail(MOLE_ERR_SHAPE, "system %u: expected 5 Jastrow parameters, or 6 with the exponent scale", s);
      if (d.n_params == (uint32_t)kSJNParam && !(d.params[5] > 0.0 && d.params[5] < 1e300))
        return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: the exponent scale must be positive and finite", s);
      if (d.n_basis < 1 || d.n_basis > 1024 || !d.basis_exponent || !d.basis_center || !d.orbital_coeff)
        return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: basis/orbital tables missing", s);
      if (d.n_orbitals < std::max(d.n_up, d.n_down))
        return fail(MOLE_ERR_SHAPE, "system %u: need n_orbitals >= max(n_up, n_down)", s);
      for (uint32_t m = 0; m < d.n_basis; ++m)
        if (d.basis_center[m] >= d.n_nuclei)
          return fail(MOLE_ERR_SHAPE, "system %u: basis function %u centred on missing nucleus", s, m);
      if (d.basis_kind)
        for (uint32_t m = 0; m < d.n_basis; ++m) {
          if (d.basis_kind[m] > MOLE_BASIS_STO_PZ)
            return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: basis function %u has unknown kind %u", s, m,
                        d.basis_kind[m]);
          if (d.basis_kind[m] == MOLE_BASIS_GTO && run->proposal == MOLE_PROPOSAL_LEARNED)
            return fail(MOLE_ERR_UNSUPPORTED, "system %u: the learned proposal takes STO bases only", s);
        }
      if (run->o_convention != MOLE_O_LOGDERIV)
        return fail(MOLE_ERR_UNSUPPORTED, "Slater-Jastrow systems report O = d ln psi / dp");
    }
    if (fam == kFamTwo && run->proposal != MOLE_PROPOSAL_BOX) {
      return fail(MOLE_ERR_UNSUPPORTED, "system %u: the two-electron kinds move with the box proposal", s);
    }
  }
  return MOLE_OK;
}

void fill_sys2(SysState &st, const mole_system_desc &d) {
  Sys2 &sp = st.sp;
  sp.kind = (int)d.wf_kind;
  sp.hamiltonian = (int)d.hamiltonian;
  sp.n_nuc = (int)d.n_nuclei;
  for (uint32_t I = 0; I < d.n_nuclei; ++I) {
    for (int k = 0; k < 3; ++k) sp.nuc[I][k] = d.nuclei_xyz[3 * I + k];
    sp.charge[I] = (double)d.nuclei_charge[I];
  }
  sp.v_nn = nuclear_repulsion(d.nuclei_xyz, d.nuclei_charge, d.n_nuclei);
  sp.zeta = d.params[0];
  sp.b = d.params[1];
  sp.c = d.params[2];
}

int default_lanes(int N) { return N <= 4 ? 4 : N <= 10 ? 8 : N <= 20 ? 16 : 32; }

mole_status check_ctx(mole_ctx *ctx, uint32_t sys) {
  if (!ctx) return fail(MOLE_ERR_INVALID_ARGUMENT, "null context");
  if (sys >= ctx->sys.size()) return fail(MOLE_ERR_INVALID_ARGUMENT, "system index %u out of range", sys);
  return MOLE_OK;
}

void free_ctx(mole_ctx *ctx) {
  if (!ctx) return;
  for (SysState &st : ctx->sys) {
    void *ptrs[] = {st.d_pos, st.d_acc, st.d_trace_mask, st.d_trace_e, st.d_samp_psi,
                    st.d_samp_g, st.d_tab, st.d_center, st.d_B, st.d_ld, st.d_zeta0};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    free_dmc(st.dmc);
    if (st.d_ops) (void)hipFree(st.d_ops);
    if (st.d_obs_acc) (void)hipFree(st.d_obs_acc);
  }
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->d_params) (void)hipFree(ctx->d_params);
  if (ctx->d_optlog) (void)hipFree(ctx->d_optlog);
  if (ctx->d_opt) (void)hipFree(ctx->d_opt);
  if (ctx->d_sysdev) (void)hipFree(ctx->d_sysdev);
  if (ctx->d_state) (void)hipFree(ctx->d_state);
  if (ctx->d_tiles) (void)hipFree(ctx->d_tiles);
  if (ctx->d_tiles_lpt) (void)hipFree(ctx->d_tiles_lpt);
  if (ctx->d_ticket) (void)hipFree(ctx->d_ticket);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// ---- Slater-Jastrow launch plumbing ----------------------------------------

int sj_tsize(const mole_ctx *ctx) { return ctx->run.precision == MOLE_PREC_MIXED ? 4 : 8; }

int sj_prop(const mole_ctx *ctx) {
  return ctx->run.proposal == MOLE_PROPOSAL_DIFFUSE   ? kSJPropDiffuse
         : ctx->run.proposal == MOLE_PROPOSAL_LEARNED ? kSJPropLearned
                                                      : kSJPropBox;
}

// Tiles of 64/L walkers grouped by LDS footprint class (4 KiB steps, the
// largest class first), most expensive first within a group
// (longest-processing-time order, so the tail of each launch is short
// tiles); each group's dynamic LDS is its largest tile footprint.
mole_status sj_build_tiles(mole_ctx *ctx) {
  struct T { SJTile t; double cost; size_t fp; };
  std::vector<T> tiles;
  size_t lds = 0;
  for (size_t s = 0; s < ctx->sys.size(); ++s) {
    SysState &st = ctx->sys[s];
    st.lds_walker = sj_lds_per_walker((int)st.N, (int)st.n_up, (int)st.n_down, st.n_nuc, st.n_basis,
                                      sj_tsize(ctx), (int)st.params.size());
    const int wpt = 64 / st.L;
    lds = std::max(lds, (size_t)wpt * st.lds_walker);
    const int nmax = (int)std::max(st.n_up, st.n_down);
    const double per_walker = (double)st.N * (st.N + st.n_nuc + st.n_basis + (double)nmax * st.n_basis +
                                               (double)nmax * nmax);
    const size_t fp = (size_t)wpt * st.lds_walker;
    for (int w0 = 0; w0 < st.W; w0 += wpt)
      tiles.push_back({{(int)s, w0, std::min(wpt, st.W - w0)}, per_walker / st.L, fp});
  }
  if (lds > ctx->lds_limit)
    return fail(MOLE_ERR_UNSUPPORTED, "walker state needs %zu B of LDS per workgroup (limit %zu): use "
                "more lanes per walker", lds, ctx->lds_limit);
  auto cls = [](size_t fp) { return (fp + 4095) / 4096; };
  std::stable_sort(tiles.begin(), tiles.end(), [&](const T &a, const T &b) {
    return cls(a.fp) != cls(b.fp) ? cls(a.fp) > cls(b.fp) : a.cost > b.cost;
  });
  std::vector<SJTile> h(tiles.size());
  ctx->tile_groups.clear();
  for (size_t i = 0; i < tiles.size(); ++i) {
    h[i] = tiles[i].t;
    if (i == 0 || cls(tiles[i].fp) != cls(tiles[i - 1].fp)) ctx->tile_groups.push_back({(int)i, 0, 0});
    mole_ctx::TileGroup &g = ctx->tile_groups.back();
    ++g.count;
    g.lds = std::max(g.lds, tiles[i].fp);
  }
  if (ctx->d_tiles) (void)hipFree(ctx->d_tiles);
  ctx->d_tiles = nullptr;
  HIP_TRY(hipMalloc(&ctx->d_tiles, sizeof(SJTile) * h.size()));
  HIP_TRY(hipMemcpy(ctx->d_tiles, h.data(), sizeof(SJTile) * h.size(), hipMemcpyHostToDevice));
  ctx->n_tiles = (int)h.size();
  ctx->lds_bytes = lds;
  // the persistent schedule's order: all tiles by cost, longest first
  std::stable_sort(tiles.begin(), tiles.end(), [](const T &a, const T &b) { return a.cost > b.cost; });
  for (size_t i = 0; i < tiles.size(); ++i) h[i] = tiles[i].t;
  if (ctx->d_tiles_lpt) (void)hipFree(ctx->d_tiles_lpt);
  ctx->d_tiles_lpt = nullptr;
  HIP_TRY(hipMalloc(&ctx->d_tiles_lpt, sizeof(SJTile) * h.size()));
  HIP_TRY(hipMemcpy(ctx->d_tiles_lpt, h.data(), sizeof(SJTile) * h.size(), hipMemcpyHostToDevice));
  if (!ctx->d_ticket) {
    HIP_TRY(hipMalloc(&ctx->d_ticket, sizeof(uint32_t)));
    HIP_TRY(hipMemset(ctx->d_ticket, 0, sizeof(uint32_t)));
  }
  const int occ = sj_persistent_occupancy(sj_tsize(ctx) == 4, lds);
  const int fill = ctx->cu_count * (occ > 0 ? occ : 1);
  ctx->persist_wg = (int)std::min<int64_t>(ctx->n_tiles, ctx->persist_wg_req ? (int64_t)ctx->persist_wg_req : fill);
  return MOLE_OK;
}

// The persistent schedule applies to plain box-proposal launches.
bool sj_persistent_eligible(const mole_ctx *ctx, const SJArgs &a) {
  return ctx->schedule == MOLE_SCHEDULE_PERSISTENT && ctx->run.proposal == MOLE_PROPOSAL_BOX && !a.ext &&
         a.dmc == kSJDmcOff;
}

// The device view of system s.
SJSysDev sj_sysdev_of(const mole_ctx *ctx, size_t s) {
  const SysState &st = ctx->sys[s];
  SJSysDev d{};
  d.N = (int)st.N;
  d.n_up = (int)st.n_up;
  d.n_dn = (int)st.n_down;
  d.n_nuc = st.n_nuc;
  d.n_basis = st.n_basis;
  d.L = st.L;
  d.hamiltonian = (int)st.hamiltonian;
  d.W = st.W;
  d.nuc = st.d_tab;
  d.charge = st.d_tab + 3 * st.n_nuc;
  d.zeta = st.d_tab + 4 * st.n_nuc;
  d.center = st.d_center;
  d.coef_t = st.d_tab + 4 * st.n_nuc + st.n_basis;
  d.n_orb = st.n_orb;
  for (int k = 0; k < kSJNParam; ++k) d.p[k] = k < (int)st.params.size() ? st.params[k] : 1.0;
  d.opt_exp = st.params.size() == (size_t)kSJNParam ? 1 : 0;
  d.v_nn = st.v_nn;
  d.pos = st.d_pos;
  d.B = st.d_B;
  d.acc = st.d_acc;
  d.ld = st.d_ld;
  d.trace_mask = st.d_trace_mask;
  d.trace_e = st.d_trace_e;
  d.samp_psi = st.d_samp_psi;
  d.samp_g = st.d_samp_g;
  d.draw_base = st.draw_base;
  d.move_base = st.move_base;
  d.sys_index = (uint32_t)s;
  d.lds_per_walker = (uint32_t)st.lds_walker;
  d.dmc = st.dmc.dmc;
  d.dmc_eref = st.dmc.st ? &st.dmc.st->eref : nullptr;
  d.dmc_pos_in = st.dmc.pos_s;
  d.dmc_B_in = st.dmc.B_s;
  d.dmc_el_in = st.dmc.el_s;
  d.dmc_w_in = st.dmc.wnew;
  d.kind = st.d_center + st.n_basis;
  d.any_gto = st.any_gto ? 1 : 0;
  d.any_p = st.any_p ? 1 : 0;
  d.ops = st.d_ops;
  d.custom_h = st.custom_h ? 1 : 0;
  d.n_obs = st.n_obs;
  d.obs_acc = st.d_obs_acc;
  return d;
}

mole_status sj_upload_sysdev(mole_ctx *ctx) {
  std::vector<SJSysDev> h(ctx->sys.size());
  for (size_t s = 0; s < ctx->sys.size(); ++s) h[s] = sj_sysdev_of(ctx, s);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(ctx->d_sysdev, h.data(), sizeof(SJSysDev) * h.size(), hipMemcpyHostToDevice));
  ctx->sysdev_dirty = false;
  return MOLE_OK;
}

// One launch per tile group, each with its group's dynamic LDS; or, on the
// persistent schedule, one launch over every tile.
hipError_t sj_launch_groups(const mole_ctx *ctx, SJArgs a) {
  const int fp32 = ctx->run.precision == MOLE_PREC_MIXED;
  if (sj_persistent_eligible(ctx, a)) {
    a.tiles = ctx->d_tiles_lpt;
    a.n_tiles = ctx->n_tiles;
    a.ticket = ctx->d_ticket;
    return launch_sj_persistent(a, ctx->persist_wg, fp32, ctx->lds_bytes, ctx->stream);
  }
  for (const mole_ctx::TileGroup &g : ctx->tile_groups) {
    a.tiles = ctx->d_tiles + g.first;
    const hipError_t e = launch_sj(a, g.count, fp32, sj_prop(ctx), g.lds, ctx->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

SJArgs sj_args(const mole_ctx *ctx) {
  SJArgs a{};
  a.systems = ctx->d_sysdev;
  a.tiles = ctx->d_tiles;
  a.seed = ctx->run.seed;
  a.epoch = ctx->epoch;
  const double lo = -0.5 * ctx->run.step, hi = 0.5 * ctx->run.step;
  a.scale = hi - lo;
  a.offset = lo - a.scale;
  a.tau = ctx->run.step;
  a.sqrt_tau = std::sqrt(ctx->run.step);
  a.block_size = ctx->run.block_size;
  a.lu_period = ctx->run.lu_period;
  a.state = ctx->d_state;
  a.cap_sweeps = (ctx->run.flags & MOLE_FLAG_TRACE) ? ctx->run.record_sweeps : 0;
  a.cap_samples = (ctx->run.flags & (MOLE_FLAG_TRACE | MOLE_FLAG_SAMPLES)) ? ctx->run.record_sweeps : 0;
  if (ctx->mlp) a.net = ctx->mlp->net();
  for (const SysState &s : ctx->sys) {
    // compat samples are written by the extended instantiation only
    if (s.any_gto || s.custom_h || s.n_obs || (ctx->run.flags & MOLE_FLAG_SAMPLES)) a.ext = 1;
    // learned proposal: from 8 electrons on, the register budget of one
    // wave per SIMD (no spills) beats two waves (FP64 atoms, 65,536 walkers:
    // Li, Be 1.4x and B..N 1.1x faster with two; O 1.08x, F and Ne 1.19x
    // faster with one)
    if (s.N >= 8) a.one_wave = 1;
  }
  return a;
}

// Writes the host's view of the launch state to the device copy.
mole_status sj_push_state(mole_ctx *ctx) {
  SJLaunchState v{};
  v.sweeps_done = ctx->sweeps_done;
  v.block_fill = ctx->block_fill;
  v.trace_sweeps = ctx->trace_sweeps;
  v.trace_samples = ctx->trace_samples;
  HIP_TRY(launch_sj_state_set(ctx->d_state, v, ctx->stream));
  return MOLE_OK;
}

// Enqueues one sweep launch of n sweeps (sampled from sweep ss on) and the
// state advance; shared by eager calls and graph capture.
mole_status sj_enqueue(mole_ctx *ctx, uint32_t n, uint32_t ss) {
  SJArgs a = sj_args(ctx);
  a.n_sweeps = n;
  a.sample_start = ss;
  HIP_TRY(sj_launch_groups(ctx, a));
  HIP_TRY(launch_sj_advance(ctx->d_state, n, n - std::min(ss, n), ctx->run.block_size, ctx->stream));
  return MOLE_OK;
}

// Host bookkeeping of n sweeps with n_samp sampled, for every system.
void account(mole_ctx *ctx, uint32_t chunk, uint32_t n_samp) {
  for (SysState &st : ctx->sys) {
    st.cols_valid = false;
    const uint64_t filled = (uint64_t)ctx->block_fill + n_samp;
    st.n_samples += (uint64_t)st.W * n_samp;
    st.n_blocks += (uint64_t)st.W * (filled / ctx->run.block_size);
    st.n_moves += (uint64_t)st.W * st.N * chunk;
  }
  if (ctx->family == kFamTwo) ctx->position += (uint64_t)chunk * ctx->sys[0].N * 4;
  else ctx->sweeps_done += chunk;
  ctx->block_fill = (uint32_t)(((uint64_t)ctx->block_fill + n_samp) % ctx->run.block_size);
  ctx->trace_sweeps += chunk;
  ctx->trace_samples += n_samp;
}

// A zero-sweep launch that draws the initial configurations (init) or
// rebuilds the inverses from the stored positions (refresh).
mole_status sj_prepare(mole_ctx *ctx, bool init) {
  if (ctx->sysdev_dirty) {
    mole_status st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  SJArgs a = sj_args(ctx);
  a.n_sweeps = 0;
  a.sample_start = 0;
  a.init = init ? 1 : 0;
  a.refresh = 1;
  for (SysState &s : ctx->sys) s.cols_valid = false;
  HIP_TRY(sj_launch_groups(ctx, a));
  return MOLE_OK;
}

// Enqueues n sweeps (sampled from sweep sample_start of the call on) with
// the host bookkeeping; counts the launches.
mole_status enqueue_sweeps(mole_ctx *ctx, uint32_t n, uint32_t sample_start, uint32_t *n_launches) {
  const double lo = -0.5 * ctx->run.step, hi = 0.5 * ctx->run.step;
  const double scale = hi - lo;
  const double offset = lo - scale;
  uint32_t launches = 0;
  uint32_t done = 0;
  while (done < n) {
    const uint32_t chunk = std::min(n - done, ctx->max_sweeps_per_launch);
    const uint32_t ss = (sample_start > done) ? std::min(sample_start - done, chunk) : 0;
    const uint32_t n_samp = chunk - ss;
    if (ctx->family == kFamSJ) {
      mole_status es = sj_enqueue(ctx, chunk, ss);
      if (es != MOLE_OK) return es;
      ++launches;
    }
    for (uint32_t s = 0; s < ctx->sys.size(); ++s) {
      SysState &st = ctx->sys[s];
      if (ctx->family == kFamTwo) {
        Sweep2Args a{};
        a.sp = st.sp;
        a.dparams = ctx->dev_params ? ctx->d_params : nullptr;
        a.W = st.W;
        a.pos = st.d_pos;
        a.acc = st.d_acc;
        a.fresh = ctx->fresh_acc ? 1 : 0;
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
        HIP_TRY(launch_sweep2(a, ctx->variant, ctx->block_threads, ctx->stream));
        ++launches;
      }
    }
    ctx->fresh_acc = false;
    account(ctx, chunk, n_samp);
    done += chunk;
  }
  *n_launches += launches;
  return MOLE_OK;
}

// Runs n sweeps; sweeps s >= sample_start of this call are sampled.
mole_status do_sweeps(mole_ctx *ctx, uint32_t n, uint32_t sample_start) {
  if (n == 0) return ok();
  if (ctx->run.proposal == MOLE_PROPOSAL_LEARNED && !ctx->mlp)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "learned proposal: attach a network (mole_mlp_attach) first");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->family == kFamSJ && ctx->sysdev_dirty) {
    mole_status st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  uint32_t launches = 0;
  mole_status st = enqueue_sweeps(ctx, n, sample_start, &launches);
  if (st != MOLE_OK) return st;
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  ctx->last_launches = launches;
  ctx->timed = true;
  return ok();
}

mole_status reduce_sys(mole_ctx *ctx, uint32_t sys, const double **cols) {
  SysState &s = ctx->sys[sys];
  if (!s.cols_valid) {
    HIP_TRY(hipSetDevice(ctx->device));
    s.cols.resize(s.n_cols);
    HIP_TRY(launch_reduce_columns(s.d_acc, s.W, s.n_cols, ctx->d_red, ctx->family == kFamSJ ? kSJMaxFirst : -1,
                                  ctx->family == kFamSJ ? kSJMaxCount : 0, ctx->stream));
    HIP_TRY(hipMemcpyAsync(s.cols.data(), ctx->d_red, sizeof(double) * s.n_cols, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    s.cols_valid = true;
  }
  *cols = s.cols.data();
  return MOLE_OK;
}

// Arguments of the tau entry points; *out the system.
mole_status tau_check(mole_ctx *ctx, uint32_t sys, size_t first, size_t n, const double *tau, size_t n_walkers,
                      SysState **out) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!tau) return fail(MOLE_ERR_INVALID_ARGUMENT, "null tau");
  if (!(ctx->run.flags & MOLE_FLAG_TRACE)) return fail(MOLE_ERR_DATA_ACCESS, "context has no trace");
  SysState &s = ctx->sys[sys];
  if (n_walkers != (size_t)s.W) return fail(MOLE_ERR_SHAPE, "expected %d walkers", s.W);
  const size_t rec = std::min((size_t)ctx->trace_samples, (size_t)ctx->run.record_sweeps);
  if (n < 2 || first > rec || n > rec - first)
    return fail(MOLE_ERR_SHAPE, "samples [%zu, %zu) outside the %zu recorded (n >= 2)", first, first + n, rec);
  *out = &s;
  return MOLE_OK;
}

}  // namespace

// ---- the allocation-free tau entry point's scratch ---------------------------
// mole_tau_int_noalloc stages the per-walker values in this file-scope device
// array instead of a buffer allocated per call: one copy per GPU, held until
// the process ends, holding the last call's values.  In host code the name
// `scratch` is a host-side stand-in, not a device address: each device's
// address comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address reaches kernels and copies.
constexpr size_t kTauScratchWalkers = (size_t)1 << 18;  // walkers (2 MiB of doubles)
static_assert(kTauScratchWalkers == MOLE_TAU_NOALLOC_MAX_WALKERS, "header and library limits differ");
static __device__ __align__(256) double scratch[kTauScratchWalkers];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// This device's address of `scratch`, looked up on its first use (the
// device must be current).
mole_status scratch_address(int device, double **out) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(MOLE_ERR_DEVICE, "device %d beyond the %d the scratch table holds", device, kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MOLE_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
  }
  *out = g_scratch_ptr[device];
  return MOLE_OK;
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
  if (name && cap) std::snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
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
  ctx->family = family_of(sys[0].wf_kind);
  ctx->sys.resize(n_sys);
  for (uint32_t s = 0; s < n_sys; ++s) {
    SysState &S = ctx->sys[s];
    const mole_system_desc &d = sys[s];
    S.kind = d.wf_kind;
    S.hamiltonian = d.hamiltonian;
    S.n_up = d.n_up;
    S.n_down = d.n_down;
    S.N = d.n_up + d.n_down;
    S.n_params = d.n_params;
    S.W = (int)d.n_walkers;
    S.params.assign(d.params, d.params + d.n_params);
    if (ctx->family == kFamTwo) {
      fill_sys2(S, d);
      S.n_cols = kAccColumns2;
    } else {
      S.n_nuc = (int)d.n_nuclei;
      S.n_basis = (int)d.n_basis;
      S.n_orb = (int)d.n_orbitals;
      S.v_nn = nuclear_repulsion(d.nuclei_xyz, d.nuclei_charge, d.n_nuclei);
      S.L = run->proposal == MOLE_PROPOSAL_LEARNED ? 4 : default_lanes((int)S.N);
      S.n_cols = sj_columns((int)d.n_params);
    }
  }
  auto bail = [&](mole_status s) {
    std::string msg = g_last_error;
    free_ctx(ctx);
    g_last_error = msg;
    return s;
  };
#define CTX_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return bail(fail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE, "%s: %s", \
                       #expr, hipGetErrorString(e_)));                                             \
  } while (0)
  CTX_TRY(hipGetDevice(&ctx->device));
  {
    hipDeviceProp_t prop;
    CTX_TRY(hipGetDeviceProperties(&prop, ctx->device));
    // opt-in LDS per workgroup (160 KiB on gfx950), granted to the
    // Slater-Jastrow kernel through hipFuncAttributeMaxDynamicSharedMemorySize
    ctx->lds_limit = std::max<size_t>(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
    ctx->cu_count = prop.multiProcessorCount;
    if (ctx->lds_limit == 0) ctx->lds_limit = 65536;
  }
  CTX_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  CTX_TRY(hipEventCreate(&ctx->ev0));
  CTX_TRY(hipEventCreate(&ctx->ev1));
  CTX_TRY(hipMalloc(&ctx->d_red, sizeof(double) * std::max<int>(kAccColumns2, kSJColumnsMax)));
  for (uint32_t s = 0; s < n_sys; ++s) {
    SysState &S = ctx->sys[s];
    const mole_system_desc &d = sys[s];
    const size_t W = (size_t)S.W;
    CTX_TRY(hipMalloc(&S.d_pos, sizeof(double) * 3 * S.N * W));
    CTX_TRY(hipMalloc(&S.d_acc, sizeof(double) * S.n_cols * W));
    CTX_TRY(launch_fill(S.d_pos, 3 * S.N * W, 0.0, ctx->stream));
    CTX_TRY(launch_fill(S.d_acc, S.n_cols * W, 0.0, ctx->stream));
    const size_t cap = run->record_sweeps;
    if (run->flags & MOLE_FLAG_TRACE) {
      CTX_TRY(hipMalloc(&S.d_trace_mask, sizeof(uint64_t) * cap * W));
      CTX_TRY(hipMalloc(&S.d_trace_e, sizeof(double) * cap * W));
      CTX_TRY(launch_fill((double *)S.d_trace_mask, cap * W, 0.0, ctx->stream));
      CTX_TRY(launch_fill(S.d_trace_e, cap * W, 0.0, ctx->stream));
    }
    if (run->flags & MOLE_FLAG_SAMPLES) {
      CTX_TRY(hipMalloc(&S.d_samp_psi, sizeof(double) * cap * W));
      CTX_TRY(hipMalloc(&S.d_samp_g, sizeof(double) * cap * W * S.n_params));
      if (!S.d_trace_e) {
        CTX_TRY(hipMalloc(&S.d_trace_e, sizeof(double) * cap * W));
        CTX_TRY(launch_fill(S.d_trace_e, cap * W, 0.0, ctx->stream));
      }
    }
    if (ctx->family == kFamSJ) {
      // tables: nuclei xyz (3 n_nuc), charges (n_nuc), zeta (n_basis), and the
      // orbital coefficients transposed to [n_basis][n_orb]
      std::vector<double> tab;
      tab.insert(tab.end(), d.nuclei_xyz, d.nuclei_xyz + 3 * d.n_nuclei);
      for (uint32_t I = 0; I < d.n_nuclei; ++I) tab.push_back((double)d.nuclei_charge[I]);
      S.zeta0.assign(d.basis_exponent, d.basis_exponent + d.n_basis);
      for (uint32_t m = 0; m < d.n_basis; ++m)
        tab.push_back(d.n_params == (uint32_t)kSJNParam ? d.params[5] * d.basis_exponent[m] : d.basis_exponent[m]);
      for (uint32_t m = 0; m < d.n_basis; ++m)
        for (uint32_t k = 0; k < d.n_orbitals; ++k) tab.push_back(d.orbital_coeff[(size_t)k * d.n_basis + m]);
      // centres, then the basis kinds (MOLE_BASIS_*)
      std::vector<int> cen(d.basis_center, d.basis_center + d.n_basis);
      for (uint32_t m = 0; m < d.n_basis; ++m) cen.push_back(d.basis_kind ? (int)d.basis_kind[m] : MOLE_BASIS_STO);
      S.any_gto = std::any_of(cen.begin() + d.n_basis, cen.end(), [](int k) { return k == MOLE_BASIS_GTO; });
      S.any_p = std::any_of(cen.begin() + d.n_basis, cen.end(), [](int k) { return k >= MOLE_BASIS_STO_PX; });
      CTX_TRY(hipMalloc(&S.d_tab, sizeof(double) * tab.size()));
      CTX_TRY(hipMemcpy(S.d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
      CTX_TRY(hipMalloc(&S.d_zeta0, sizeof(double) * d.n_basis));
      CTX_TRY(hipMemcpy(S.d_zeta0, d.basis_exponent, sizeof(double) * d.n_basis, hipMemcpyHostToDevice));
      CTX_TRY(hipMalloc(&S.d_center, sizeof(int) * cen.size()));
      CTX_TRY(hipMemcpy(S.d_center, cen.data(), sizeof(int) * cen.size(), hipMemcpyHostToDevice));
      const size_t nB = (size_t)S.n_up * S.n_up + (size_t)S.n_down * S.n_down;
      CTX_TRY(hipMalloc(&S.d_B, (size_t)sj_tsize(ctx) * nB * W + 8));
      CTX_TRY(launch_fill((double *)S.d_B, (sj_tsize(ctx) * nB * W) / 8, 0.0, ctx->stream));
      CTX_TRY(hipMalloc(&S.d_ld, sizeof(double) * kSJLdColumns * W));
      CTX_TRY(launch_fill(S.d_ld, kSJLdColumns * W, 0.0, ctx->stream));
    }
  }
  if (ctx->family == kFamSJ) {
    CTX_TRY(hipMalloc(&ctx->d_sysdev, sizeof(SJSysDev) * n_sys));
    CTX_TRY(hipMalloc(&ctx->d_state, sizeof(SJLaunchState)));
    if (sj_push_state(ctx) != MOLE_OK) return bail(MOLE_ERR_DEVICE);
    mole_status ts = sj_build_tiles(ctx);
    if (ts != MOLE_OK) return bail(ts);
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
  if (ctx->family == kFamTwo) {
    for (uint32_t s = 0; s < ctx->sys.size(); ++s) {
      SysState &ss = ctx->sys[s];
      HIP_TRY(launch_init2(ss.d_pos, ss.W, ctx->run.seed, s, ctx->epoch, ctx->stream));
    }
    ctx->position = 3ull * ctx->sys[0].N;
    return ok();
  }
  for (SysState &s : ctx->sys) {
    s.draw_base = 3ull * s.N;  // moves start after Sampler::new's 3N draws
    s.move_base = 0;
  }
  ctx->sweeps_done = 0;
  ctx->sysdev_dirty = true;
  st = sj_push_state(ctx);
  if (st != MOLE_OK) return st;
  st = sj_prepare(ctx, true);
  return st == MOLE_OK ? ok() : st;
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
  if (ctx->family == kFamSJ) {
    st = sj_prepare(ctx, false);
    if (st != MOLE_OK) return st;
  }
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
  if (ctx->family == kFamSJ) {
    st = sj_prepare(ctx, false);
    if (st != MOLE_OK) return st;
  }
  return ok();
}

mole_status mole_set_stream(mole_ctx *ctx, uint32_t epoch, uint64_t position) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  ctx->epoch = epoch;
  ctx->position = position;
  if (ctx->family == kFamSJ) {
    for (SysState &s : ctx->sys) {
      s.draw_base = position;
      s.move_base = 0;
    }
    ctx->sweeps_done = 0;
    ctx->sysdev_dirty = true;
    HIP_TRY(hipSetDevice(ctx->device));
    st = sj_push_state(ctx);
    if (st != MOLE_OK) return st;
  }
  return ok();
}

mole_status mole_get_stream(mole_ctx *ctx, uint32_t *epoch, uint64_t *position) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (epoch) *epoch = ctx->epoch;
  if (position)
    *position = ctx->family == kFamTwo
                    ? ctx->position
                    : ctx->sys[0].draw_base + (uint64_t)sj_draws_per_move(sj_prop(ctx)) *
                                                    ctx->sys[0].N * ctx->sweeps_done;
  return ok();
}

mole_status mole_set_params(mole_ctx *ctx, uint32_t sys, const double *p, uint32_t n_params) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!p || n_params != s.n_params) return fail(MOLE_ERR_SHAPE, "expected %u parameters", s.n_params);
  if (ctx->family == kFamSJ && n_params == (uint32_t)kSJNParam) {
    if (!(p[5] > 0.0 && p[5] < 1e300)) return fail(MOLE_ERR_INVALID_ARGUMENT, "the exponent scale must be positive");
    // the exponents the kernels read: s zeta0_m, after the queued sweeps
    std::vector<double> z(s.zeta0.size());
    for (size_t m = 0; m < z.size(); ++m) z[m] = p[5] * s.zeta0[m];
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(s.d_tab + 4 * s.n_nuc, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
  }
  s.params.assign(p, p + n_params);
  if (ctx->family == kFamTwo) {
    s.sp.zeta = p[0];
    s.sp.b = p[1];
    s.sp.c = p[2];
  } else {
    ctx->sysdev_dirty = true;
  }
  return ok();
}

mole_status mole_get_params(mole_ctx *ctx, uint32_t sys, double *p, uint32_t n_params) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!p || n_params != s.n_params) return fail(MOLE_ERR_SHAPE, "expected %u parameters", s.n_params);
  std::copy(s.params.begin(), s.params.end(), p);
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
    for (SysState &s : ctx->sys) {
      s.cols_valid = false;
      const int col = ctx->family == kFamTwo ? (int)kAccBlockSum : (int)kSJBlockSum;
      HIP_TRY(launch_fill(s.d_acc + (size_t)col * s.W, (size_t)s.W, 0.0, ctx->stream));
    }
    ctx->block_fill = 0;
    if (ctx->family == kFamSJ) {
      st = sj_push_state(ctx);
      if (st != MOLE_OK) return st;
    }
  }
  const uint32_t blocks = steps / block_size;
  return do_sweeps(ctx, blocks * block_size, block_size);
}

mole_status mole_reset_stats(mole_ctx *ctx) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  for (SysState &s : ctx->sys) {
    HIP_TRY(launch_fill(s.d_acc, (size_t)s.n_cols * s.W, 0.0, ctx->stream));
    if (s.d_obs_acc) HIP_TRY(launch_fill(s.d_obs_acc, 2 * (size_t)MOLE_MAX_OBSERVABLES * s.W, 0.0, ctx->stream));
    s.n_samples = s.n_blocks = s.n_moves = 0;
    s.cols_valid = false;
  }
  ctx->block_fill = 0;
  ctx->trace_sweeps = ctx->trace_samples = 0;
  if (ctx->family == kFamSJ) {
    st = sj_push_state(ctx);
    if (st != MOLE_OK) return st;
  }
  return ok();
}

mole_status mole_read_stats(mole_ctx *ctx, uint32_t sys, mole_stats *out) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  const double *cols = nullptr;
  st = reduce_sys(ctx, sys, &cols);
  if (st != MOLE_OK) return st;
  const SysState &s = ctx->sys[sys];
  const bool sj = ctx->family == kFamSJ;
  out->n_samples = s.n_samples;
  out->n_blocks = s.n_blocks;
  out->n_moves = s.n_moves;
  out->n_accepted = (uint64_t)cols[sj ? kSJAccepted : kAccAccepted];
  out->sum_e = cols[sj ? kSJSumE : kAccSumE];
  out->sum_e2 = cols[sj ? kSJSumE2 : kAccSumE2];
  out->sum_block_mean = cols[sj ? kSJSumBm : kAccSumBm];
  out->sum_block_mean2 = cols[sj ? kSJSumBm2 : kAccSumBm2];
  out->max_condition = sj ? cols[kSJCondMax] : 0.0;
  out->n_singular = sj ? (uint64_t)cols[kSJSingular] : 0;
  out->max_logdet_drift = sj ? cols[kSJLdDrift] : 0.0;
  return ok();
}

mole_status mole_read_sr_sums(mole_ctx *ctx, uint32_t sys, mole_sr_sums *out) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  const double *cols = nullptr;
  st = reduce_sys(ctx, sys, &cols);
  if (st != MOLE_OK) return st;
  std::memset(out, 0, sizeof *out);
  const bool sj = ctx->family == kFamSJ;
  // the system's parameters (a Slater-Jastrow system accumulates its own 5 or 6)
  const int np = sj ? (int)ctx->sys[sys].n_params : kNParam2;
  const int nacc = np;  // the accumulator triangle's order
  const int cO = sj ? kSJSumO : kAccSumO, cOE = sj ? sj_sum_oe(np) : kAccSumOE, cOO = sj ? sj_sum_oo(np) : kAccSumOO;
  out->n_params = np;
  out->n = ctx->sys[sys].n_samples;
  out->sum_e = cols[sj ? kSJSumE : kAccSumE];
  for (int k = 0; k < np; ++k) {
    out->sum_o[k] = cols[cO + k];
    out->sum_oe[k] = cols[cOE + k];
    for (int l = k; l < np; ++l) {
      const int t = k * nacc - k * (k - 1) / 2 + (l - k);  // upper-triangle position of (k, l)
      out->sum_oo[k * np + l] = cols[cOO + t];
      out->sum_oo[l * np + k] = cols[cOO + t];
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

mole_status mole_read_blocks(mole_ctx *ctx, uint32_t sys, size_t first, mole_block_record *out, size_t cap,
                             size_t *n) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!n) return fail(MOLE_ERR_INVALID_ARGUMENT, "null count");
  if (!(ctx->run.flags & MOLE_FLAG_TRACE)) return fail(MOLE_ERR_DATA_ACCESS, "context has no trace");
  SysState &s = ctx->sys[sys];
  const size_t B = ctx->run.block_size, rec = ctx->run.record_sweeps;
  const size_t samples = std::min((size_t)ctx->trace_samples, rec);
  const size_t complete = samples / B;
  const size_t k = first < complete ? std::min(cap, complete - first) : 0;
  *n = out ? k : (first < complete ? complete - first : 0);
  if (!out || k == 0) return ok();
  HIP_TRY(hipSetDevice(ctx->device));
  double *d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * k));
  hipError_t e = launch_block_records(s.d_trace_e, s.W, rec, B, first, k, d, ctx->stream);
  std::vector<double> h(2 * k);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_read_blocks: %s", hipGetErrorString(e));
  for (size_t i = 0; i < k; ++i) {
    out[i].block = first + i;
    out[i].n_walkers = (uint64_t)s.W;
    out[i].sum_block_mean = h[2 * i];
    out[i].sum_block_mean2 = h[2 * i + 1];
    out[i].mean = h[2 * i] / (double)s.W;
  }
  return ok();
}

mole_status mole_tau_int(mole_ctx *ctx, uint32_t sys, size_t first, size_t n, double *tau, size_t n_walkers) {
  SysState *sp = nullptr;
  mole_status st = tau_check(ctx, sys, first, n, tau, n_walkers, &sp);
  if (st != MOLE_OK) return st;
  SysState &s = *sp;
  HIP_TRY(hipSetDevice(ctx->device));
  double *d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(double) * (size_t)s.W));
  hipError_t e = launch_tau_corr(s.d_trace_e, s.W, ctx->run.record_sweeps, first, (int)n, d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tau, d, sizeof(double) * (size_t)s.W, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_tau_int: %s", hipGetErrorString(e));
  return ok();
}

mole_status mole_tau_int_noalloc(mole_ctx *ctx, uint32_t sys, size_t first, size_t n, double *tau,
                                 size_t n_walkers) {
  SysState *sp = nullptr;
  mole_status st = tau_check(ctx, sys, first, n, tau, n_walkers, &sp);
  if (st != MOLE_OK) return st;
  SysState &s = *sp;
  if ((size_t)s.W > kTauScratchWalkers)
    return fail(MOLE_ERR_UNSUPPORTED, "mole_tau_int_noalloc: %d walkers exceed its scratch of %zu (use mole_tau_int)",
                s.W, kTauScratchWalkers);
  HIP_TRY(hipSetDevice(ctx->device));
  double *d = nullptr;
  st = scratch_address(ctx->device, &d);
  if (st != MOLE_OK) return st;
  hipError_t e = launch_tau_corr(s.d_trace_e, s.W, ctx->run.record_sweeps, first, (int)n, d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tau, d, sizeof(double) * (size_t)s.W, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_tau_int_noalloc: %s", hipGetErrorString(e));
  return ok();
}

mole_status mole_read_samples(mole_ctx *ctx, uint32_t sys, double *e_local, double *psi, double *pgrad,
                              size_t walkers, size_t samples) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (!(ctx->run.flags & MOLE_FLAG_SAMPLES)) return fail(MOLE_ERR_DATA_ACCESS, "context keeps no samples");
  const size_t cap = ctx->run.record_sweeps;
  if (walkers != (size_t)s.W || samples > cap)
    return fail(MOLE_ERR_SHAPE, "sample request exceeds %d walkers x %zu records", s.W, cap);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const size_t np = s.n_params;  // parameters per "Parameter gradient" sample
  std::vector<double> h(cap * s.W * np);
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
    HIP_TRY(hipMemcpy(h.data(), s.d_samp_g, cap * s.W * np * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < walkers; ++w)
      std::memcpy(pgrad + w * samples * np, h.data() + w * cap * np, samples * np * sizeof(double));
  }
  return ok();
}

mole_status mole_read_logdet(mole_ctx *ctx, uint32_t sys, double *logdet, double *sign, size_t n_walkers) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "log-determinants are kept for Slater-Jastrow systems");
  if (!logdet) return fail(MOLE_ERR_INVALID_ARGUMENT, "null logdet");
  SysState &s = ctx->sys[sys];
  if (n_walkers != (size_t)s.W) return fail(MOLE_ERR_SHAPE, "expected %d walkers", s.W);
  const size_t W = (size_t)s.W;
  std::vector<double> h((size_t)kSJLdColumns * W);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(h.data(), s.d_ld, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (size_t w = 0; w < W; ++w)
    for (int sp = 0; sp < 2; ++sp) {
      const double mant = h[(size_t)(2 * sp) * W + w], ex = h[(size_t)(2 * sp + 1) * W + w];
      // det = mant 2^ex
      logdet[2 * w + sp] = mant == 0.0 ? -INFINITY : std::log(std::fabs(mant)) + ex * 0.69314718055994530942;
      if (sign) sign[2 * w + sp] = mant > 0.0 ? 1.0 : (mant < 0.0 ? -1.0 : 0.0);
    }
  return ok();
}

// mole_eval of a Slater-Jastrow system: the configurations become the
// walkers of a temporary copy of the system (positions [3N][n], inverses,
// accumulators and determinants of its own), evaluated by one launch of the
// extended kernel in its evaluation mode (re-inversion, then the local
// energy); the context's walkers are not touched.
static mole_status sj_eval(mole_ctx *ctx, uint32_t sys, size_t n, const double *cfg, double *value, double *grad,
                           double *lap, double *e_local, double *pgrad) {
  const SysState &s = ctx->sys[sys];
  const size_t D = 3 * (size_t)s.N, np = s.params.size();
  const size_t nB = (size_t)s.n_up * s.n_up + (size_t)s.n_down * s.n_down;
  const size_t ncol = (size_t)sj_columns((int)np);
  const int wpt = 64 / s.L;
  if (n > ((size_t)1 << 28)) return fail(MOLE_ERR_SHAPE, "mole_eval: at most 2^28 configurations per call");
  // evaluated in FP64 whatever the context's precision (Function::value and
  // its derivatives at given configurations, no update chain to save)
  const size_t lds_walker = sj_lds_per_walker((int)s.N, (int)s.n_up, (int)s.n_down, s.n_nuc, s.n_basis, 8, (int)np);
  if ((size_t)wpt * lds_walker > ctx->lds_limit)
    return fail(MOLE_ERR_UNSUPPORTED, "mole_eval: %zu B of LDS per workgroup at %d lanes per walker (limit %zu)",
                (size_t)wpt * lds_walker, s.L, ctx->lds_limit);
  const size_t n_tiles = (n + wpt - 1) / wpt;
  // one allocation of doubles: positions, inverses (8 bytes per element for
  // either precision), accumulators, determinants, then the outputs
  const size_t o_pos = 0, o_B = o_pos + D * n, o_acc = o_B + nB * n, o_ld = o_acc + ncol * n,
               o_val = o_ld + kSJLdColumns * n, o_grad = o_val + n, o_lap = o_grad + D * n, o_el = o_lap + n,
               o_pg = o_el + n, total = o_pg + np * n;
  double *d = nullptr;
  SJSysDev *d_sys = nullptr;
  SJTile *d_tiles = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(&d, sizeof(double) * total);
  if (e == hipSuccess) e = hipMalloc(&d_sys, sizeof(SJSysDev));
  if (e == hipSuccess) e = hipMalloc(&d_tiles, sizeof(SJTile) * n_tiles);
  if (e == hipSuccess) {
    std::vector<double> soa(D * n);
    for (size_t w = 0; w < n; ++w)
      for (size_t k = 0; k < D; ++k) soa[k * n + w] = cfg[w * D + k];
    SJSysDev v = sj_sysdev_of(ctx, sys);
    v.W = (int)n;
    v.lds_per_walker = (uint32_t)lds_walker;
    v.pos = d + o_pos;
    v.B = d + o_B;
    v.acc = d + o_acc;
    v.ld = d + o_ld;
    v.trace_mask = nullptr;
    v.trace_e = nullptr;
    v.samp_psi = v.samp_g = nullptr;
    v.dmc = nullptr;
    v.dmc_eref = nullptr;
    v.dmc_pos_in = nullptr;
    v.dmc_B_in = nullptr;
    v.dmc_el_in = v.dmc_w_in = nullptr;
    v.n_obs = 0;
    v.obs_acc = nullptr;
    v.ev_value = value ? d + o_val : nullptr;
    v.ev_grad = grad ? d + o_grad : nullptr;
    v.ev_lap = lap ? d + o_lap : nullptr;
    v.ev_el = e_local ? d + o_el : nullptr;
    v.ev_pgrad = pgrad ? d + o_pg : nullptr;
    std::vector<SJTile> tiles(n_tiles);
    for (size_t k = 0; k < n_tiles; ++k)
      tiles[k] = SJTile{0, (int)(k * wpt), (int)std::min<size_t>(wpt, n - k * wpt)};
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(d_sys, &v, sizeof(SJSysDev), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tiles, tiles.data(), sizeof(SJTile) * n_tiles, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_fill(d + o_B, o_val - o_B, 0.0, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d + o_pos, soa.data(), sizeof(double) * D * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      SJArgs a = sj_args(ctx);
      a.systems = d_sys;
      a.tiles = d_tiles;
      a.n_sweeps = 0;
      a.sample_start = 0;
      a.init = 0;
      a.refresh = 1;
      a.dmc = kSJEval;
      a.ext = 1;
      a.one_wave = 0;
      a.cap_sweeps = a.cap_samples = 0;
      e = launch_sj(a, (int)n_tiles, 0, kSJPropBox, (size_t)wpt * lds_walker, ctx->stream);
    }
    auto back = [&](double *host, size_t off, size_t cnt) {
      if (e == hipSuccess && host)
        e = hipMemcpyAsync(host, d + off, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream);
    };
    back(value, o_val, n);
    back(grad, o_grad, D * n);
    back(lap, o_lap, n);
    back(e_local, o_el, n);
    back(pgrad, o_pg, np * n);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  (void)hipFree(d_tiles);
  (void)hipFree(d_sys);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_eval: %s", hipGetErrorString(e));
  return ok();
}

mole_status mole_eval(mole_ctx *ctx, uint32_t sys, size_t n, const double *cfg, double *value,
                      double *grad, double *lap, double *e_local, double *pgrad) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null cfg");
  if (n == 0) return ok();
  if (ctx->family == kFamSJ) return sj_eval(ctx, sys, n, cfg, value, grad, lap, e_local, pgrad);
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
    if (e == hipSuccess && host) e = hipMemcpyAsync(host, dev, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream);
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

mole_status mole_compute_parameter_update(const mole_sr_sums *sums, uint32_t optimizer, double step,
                                          uint32_t sr_mode, double sr_eps, double *dp, double *g_out) {
  if (!sums || !dp) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  const int np = (int)sums->n_params;
  if (np < 1 || np > MOLE_MAX_PARAMS) return fail(MOLE_ERR_SHAPE, "n_params out of range");
  if (sums->n == 0) return fail(MOLE_ERR_DATA_ACCESS, "no samples");
  if (optimizer > MOLE_OPT_SR || sr_mode > MOLE_SR_REFERENCE)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown optimizer or SR mode");
  const int rc = sr_update(np, (double)sums->n, sums->sum_e, sums->sum_o, sums->sum_oe, sums->sum_oo, optimizer, step,
                           sr_mode, sr_eps, dp, g_out);
  if (rc == MOLE_ERR_LINALG) return fail(MOLE_ERR_LINALG, "SR matrix is singular");
  return ok();
}

mole_status mole_optimizer_init(mole_optimizer *opt, uint32_t kind, uint32_t n_params, double step, double momentum,
                                uint32_t history, uint32_t sr_mode, double sr_eps) {
  if (!opt) return fail(MOLE_ERR_INVALID_ARGUMENT, "null optimizer");
  if (kind > MOLE_OPT_LBFGS) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown optimizer %u", kind);
  if (n_params < 1 || n_params > MOLE_MAX_PARAMS) return fail(MOLE_ERR_SHAPE, "n_params out of range");
  if (kind == MOLE_OPT_SR && sr_mode > MOLE_SR_REFERENCE) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown SR mode");
  if (kind == MOLE_OPT_LBFGS && (history < 1 || history > MOLE_LBFGS_MAX_HISTORY))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "LBFGS history must be 1..%d", MOLE_LBFGS_MAX_HISTORY);
  std::memset(opt, 0, sizeof(*opt));
  opt->kind = kind;
  opt->n_params = n_params;
  opt->sr_mode = kind == MOLE_OPT_SR ? sr_mode : 0;
  opt->history = kind == MOLE_OPT_LBFGS ? history : 0;
  opt->step = step;
  opt->sr_eps = sr_eps;
  opt->momentum = momentum;
  if (kind == MOLE_OPT_LBFGS)
    for (uint32_t i = 0; i < n_params; ++i) opt->grad_prev[i] = opt->pars_prev[i] = 1e-5;  // optimizers.rs:109,115-116
  return ok();
}

static mole_status check_optimizer(const mole_optimizer *opt) {
  if (!opt) return fail(MOLE_ERR_INVALID_ARGUMENT, "null optimizer");
  if (opt->kind > MOLE_OPT_LBFGS || opt->n_params < 1 || opt->n_params > MOLE_MAX_PARAMS ||
      (opt->kind == MOLE_OPT_SR && opt->sr_mode > MOLE_SR_REFERENCE) ||
      (opt->kind == MOLE_OPT_LBFGS && (opt->history < 1 || opt->history > MOLE_LBFGS_MAX_HISTORY ||
                                       opt->n_pairs > opt->history || opt->head >= opt->history)))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "optimizer state is not one mole_optimizer_init made");
  return ok();
}

mole_status mole_optimizer_update(mole_optimizer *opt, const mole_sr_sums *sums, const double *pars, double *dp,
                                  double *g) {
  mole_status st = check_optimizer(opt);
  if (st != MOLE_OK) return st;
  if (!sums || !dp) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  if (sums->n_params != opt->n_params) return fail(MOLE_ERR_SHAPE, "sums hold %u parameters, the optimizer %u",
                                                   sums->n_params, opt->n_params);
  if (sums->n == 0) return fail(MOLE_ERR_DATA_ACCESS, "no samples");
  if (opt->kind == MOLE_OPT_LBFGS && !pars) return fail(MOLE_ERR_INVALID_ARGUMENT, "LBFGS needs the parameters");
  const int rc = optimizer_update(opt, (double)sums->n, sums->sum_e, sums->sum_o, sums->sum_oe, sums->sum_oo, pars,
                                  dp, g);
  if (rc == MOLE_ERR_LINALG) return fail(MOLE_ERR_LINALG, "SR matrix is singular");
  return ok();
}

mole_status mole_set_launch_config(mole_ctx *ctx, uint32_t variant, uint32_t block_threads) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (variant > MOLE_VARIANT_QUAD) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown variant %u", variant);
  if (block_threads != 0 && (block_threads % 64 != 0 || block_threads > 256))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "block_threads must be 0 or a multiple of 64 up to 256");
  if (ctx->family == kFamTwo && (variant == MOLE_VARIANT_WAVESPEC || variant == MOLE_VARIANT_QUAD))
    for (const SysState &s : ctx->sys)
      if (!sweep2_wavespec_eligible(s.sp))
        return fail(MOLE_ERR_UNSUPPORTED, "the wave-specialised and quad kernels need <= 2 nuclei (1 for helium)");
  ctx->variant = (int)variant;
  if (block_threads) ctx->block_threads = (int)block_threads;
  return ok();
}

mole_status mole_set_walker_lanes(mole_ctx *ctx, uint32_t sys, uint32_t lanes) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "lanes per walker apply to Slater-Jastrow systems");
  if (lanes > 64 || (lanes & (lanes - 1)) != 0)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "lanes must be a power of two up to 64 (or 0)");
  if (ctx->run.proposal == MOLE_PROPOSAL_LEARNED && lanes != 0 && lanes != 4)
    return fail(MOLE_ERR_UNSUPPORTED, "the learned proposal runs 16 walkers of 4 lanes per wave");
  SysState &s = ctx->sys[sys];
  const int old = s.L;
  s.L = lanes ? (int)lanes : (ctx->run.proposal == MOLE_PROPOSAL_LEARNED ? 4 : default_lanes((int)s.N));
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  st = sj_build_tiles(ctx);
  if (st != MOLE_OK) {
    s.L = old;
    (void)sj_build_tiles(ctx);
    return st;
  }
  ctx->sysdev_dirty = true;
  ++ctx->generation;
  return ok();
}

mole_status mole_get_walker_lanes(mole_ctx *ctx, uint32_t sys, uint32_t *lanes) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!lanes) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  *lanes = (uint32_t)ctx->sys[sys].L;
  return ok();
}

struct mole_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t n_sweeps = 0, n_samp = 0, generation = 0;
  const mole_ctx *ctx = nullptr;
};

void mole_graph_destroy(mole_graph *g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

mole_status mole_graph_capture(mole_ctx *ctx, uint32_t n_sweeps, uint32_t flags, mole_graph **out) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "graph capture covers Slater-Jastrow contexts");
  if (n_sweeps == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "n_sweeps must be >= 1");
  if (ctx->run.proposal == MOLE_PROPOSAL_LEARNED && !ctx->mlp)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "learned proposal: attach a network (mole_mlp_attach) first");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->sysdev_dirty) {
    st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  const uint32_t ss = (flags & MOLE_SWEEP_SAMPLE) ? 0u : n_sweeps;
  mole_graph *g = new mole_graph();
  g->n_sweeps = n_sweeps;
  g->n_samp = n_sweeps - ss;
  g->generation = ctx->generation;
  g->ctx = ctx;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    delete g;
    return fail(MOLE_ERR_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  st = sj_enqueue(ctx, n_sweeps, ss);
  e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (st != MOLE_OK || e != hipSuccess) {
    mole_graph_destroy(g);
    return st != MOLE_OK ? st : fail(MOLE_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    mole_graph_destroy(g);
    return fail(MOLE_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *out = g;
  return ok();
}

mole_status mole_graph_launch(mole_ctx *ctx, mole_graph *g, uint32_t replays) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (!g || g->ctx != ctx) return fail(MOLE_ERR_INVALID_ARGUMENT, "graph does not belong to this context");
  if (g->generation != ctx->generation)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "graph is stale: the launch shape changed after capture");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->sysdev_dirty) {
    st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  for (uint32_t r = 0; r < replays; ++r) {
    HIP_TRY(hipGraphLaunch(g->exec, ctx->stream));
    account(ctx, g->n_sweeps, g->n_samp);
  }
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  ctx->last_launches = replays;
  ctx->timed = true;
  return ok();
}

mole_status mole_get_launch_info(mole_ctx *ctx, uint32_t *workgroups, size_t *lds_bytes) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "launch shape applies to Slater-Jastrow contexts");
  const bool persist = ctx->schedule == MOLE_SCHEDULE_PERSISTENT && ctx->run.proposal == MOLE_PROPOSAL_BOX;
  if (workgroups) *workgroups = (uint32_t)(persist ? ctx->persist_wg : ctx->n_tiles);
  if (lds_bytes) *lds_bytes = ctx->lds_bytes;
  return ok();
}

mole_status mole_set_tile_schedule(mole_ctx *ctx, uint32_t schedule, uint32_t workgroups) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "tile schedules apply to Slater-Jastrow contexts");
  if (schedule > MOLE_SCHEDULE_PERSISTENT) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown schedule %u", schedule);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->schedule = (int)schedule;
  ctx->persist_wg_req = workgroups;
  st = sj_build_tiles(ctx);
  if (st != MOLE_OK) return st;
  ++ctx->generation;
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

// ---- learned proposal: the network, its data and its training -------------

namespace {

void free_mlp(mole_mlp *m) {
  if (!m) return;
  void *ptrs[] = {m->d_w, m->d_m, m->d_v, m->d_grad, m->d_fwd1, m->d_fwd, m->d_bwd, m->d_bias,
                  m->d_cfg, m->d_partial, m->d_obj, m->d_obj_steps, m->d_mid, m->d_mid_obj};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

mole_status mlp_pack(mole_mlp *m) {
  HIP_TRY(launch_mlp_pack(m->d_w, m->d_fwd1, m->d_fwd, m->d_bwd, m->d_bias, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return MOLE_OK;
}

// The training kernel keeps a sample's configuration and one spin block in
// fixed-size LDS arrays.
mole_status mlp_check_system(const SysState &s) {
  if (s.N > 16 || s.n_up > 8 || s.n_down > 8)
    return fail(MOLE_ERR_UNSUPPORTED, "learned-proposal training supports N <= 16 and spin blocks <= 8");
  if (s.any_gto) return fail(MOLE_ERR_UNSUPPORTED, "learned-proposal training takes STO bases only");
  return MOLE_OK;
}

}  // namespace

mole_status mole_mlp_create(uint64_t seed, mole_mlp **out) {
  if (!out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  mole_mlp *m = new mole_mlp();
  auto bail = [&](mole_status s) {
    std::string msg = g_last_error;
    free_mlp(m);
    g_last_error = msg;
    return s;
  };
#define MLP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      fail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE, "%s: %s", #expr, \
           hipGetErrorString(e_));                                                          \
      return bail(e_ == hipErrorOutOfMemory ? MOLE_ERR_OUT_OF_MEMORY : MOLE_ERR_DEVICE);    \
    }                                                                                       \
  } while (0)
  MLP_TRY(hipGetDevice(&m->device));
  MLP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MLP_TRY(hipMalloc(&m->d_w, sizeof(float) * kMlpParams));
  MLP_TRY(hipMalloc(&m->d_m, sizeof(float) * kMlpParams));
  MLP_TRY(hipMalloc(&m->d_v, sizeof(float) * kMlpParams));
  MLP_TRY(hipMalloc(&m->d_grad, sizeof(float) * kMlpParams));
  MLP_TRY(hipMalloc(&m->d_fwd1, sizeof(bf16x4) * kFwd1Frags * 64));
  MLP_TRY(hipMalloc(&m->d_fwd, sizeof(bf16x8) * kFwdFrags * 64));
  MLP_TRY(hipMalloc(&m->d_bwd, sizeof(bf16x4) * kBwdFrags * 64));
  MLP_TRY(hipMalloc(&m->d_bias, sizeof(float) * kMlpBiasPad));
  MLP_TRY(launch_fill((double *)m->d_m, kMlpParams / 2, 0.0, m->stream));
  MLP_TRY(launch_fill((double *)m->d_v, kMlpParams / 2, 0.0, m->stream));
  MLP_TRY(launch_mlp_init(m->d_w, seed, m->stream));
#undef MLP_TRY
  mole_status st = mlp_pack(m);
  if (st != MOLE_OK) return bail(st);
  *out = m;
  return ok();
}

void mole_mlp_destroy(mole_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_mlp(m);
}

mole_status mole_mlp_set_weights(mole_mlp *m, const float *w, size_t n) {
  if (!m || !w) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  if (n != (size_t)kMlpParams) return fail(MOLE_ERR_SHAPE, "expected %d parameters", kMlpParams);
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpyAsync(m->d_w, w, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(launch_fill((double *)m->d_m, kMlpParams / 2, 0.0, m->stream));
  HIP_TRY(launch_fill((double *)m->d_v, kMlpParams / 2, 0.0, m->stream));
  m->adam_step = 0;
  mole_status st = mlp_pack(m);
  return st == MOLE_OK ? ok() : st;
}

mole_status mole_mlp_get_weights(mole_mlp *m, float *w, size_t n) {
  if (!m || !w) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  if (n != (size_t)kMlpParams) return fail(MOLE_ERR_SHAPE, "expected %d parameters", kMlpParams);
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpyAsync(w, m->d_w, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return ok();
}

mole_status mole_mlp_attach(mole_ctx *ctx, mole_mlp *m) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "the learned proposal moves Slater-Jastrow systems");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->mlp = m;
  ++ctx->generation;  // captured graphs hold the old network's addresses
  return ok();
}

mole_status mole_mlp_collect(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t n_snapshots,
                             uint32_t sweeps_between) {
  if (!m) return fail(MOLE_ERR_INVALID_ARGUMENT, "null network");
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "training data come from Slater-Jastrow systems");
  SysState &s = ctx->sys[sys];
  st = mlp_check_system(s);
  if (st != MOLE_OK) return st;
  if (m->n_cfg && m->data_N != s.N)
    return fail(MOLE_ERR_SHAPE, "the training set holds %u-electron configurations", m->data_N);
  if (sweeps_between == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "sweeps_between must be >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t need = m->n_cfg + (size_t)n_snapshots * s.W;
  if (need > m->cap_cfg) {
    size_t cap = std::max(need, 2 * m->cap_cfg);
    double *nb = nullptr;
    HIP_TRY(hipMalloc(&nb, sizeof(double) * 3 * s.N * cap));
    if (m->d_cfg) {
      HIP_TRY(hipMemcpyAsync(nb, m->d_cfg, sizeof(double) * 3 * s.N * m->n_cfg, hipMemcpyDeviceToDevice,
                             m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      (void)hipFree(m->d_cfg);
    }
    m->d_cfg = nb;
    m->cap_cfg = cap;
  }
  m->data_N = s.N;
  for (uint32_t k = 0; k < n_snapshots; ++k) {
    st = do_sweeps(ctx, sweeps_between, sweeps_between);
    if (st != MOLE_OK) return st;
    // [3N][W] -> [W][3N]: a strided 2D copy
    for (uint32_t c = 0; c < 3 * s.N; ++c)
      HIP_TRY(hipMemcpy2DAsync(m->d_cfg + m->n_cfg * 3 * s.N + c, sizeof(double) * 3 * s.N,
                               s.d_pos + (size_t)c * s.W, sizeof(double), sizeof(double), (size_t)s.W,
                               hipMemcpyDeviceToDevice, ctx->stream));
    m->n_cfg += s.W;
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ok();
}

mole_status mole_mlp_clear_data(mole_mlp *m) {
  if (!m) return fail(MOLE_ERR_INVALID_ARGUMENT, "null network");
  m->n_cfg = 0;
  m->data_N = 0;
  return ok();
}

mole_status mole_mlp_data_size(mole_mlp *m, size_t *n_cfg) {
  if (!m || !n_cfg) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  *n_cfg = m->n_cfg;
  return ok();
}

namespace {

// One gradient evaluation of `batch` samples starting at global sample
// `first` into m->d_grad and *obj_dev (device); returns after enqueueing.
mole_status mlp_gradient_enqueue(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t batch, uint64_t first,
                                 uint64_t seed, double *obj_dev, const MlpAdamStep *adam = nullptr,
                                 uint32_t objective = MOLE_MLP_OBJ_ESJD) {
  const int n_mb = (int)(batch / 16);
  const int G = std::min(n_mb, 1024);
  if (G > m->partial_groups) {
    for (void *p : {(void *)m->d_partial, (void *)m->d_obj, (void *)m->d_mid, (void *)m->d_mid_obj})
      if (p) (void)hipFree(p);
    m->d_partial = m->d_mid = nullptr;
    m->d_obj = m->d_mid_obj = nullptr;
    m->partial_groups = 0;
    const size_t H = ((size_t)G + kMlpSlice - 1) / kMlpSlice;
    HIP_TRY(hipMalloc(&m->d_partial, sizeof(float) * kMlpParams * (size_t)G));
    HIP_TRY(hipMalloc(&m->d_obj, sizeof(double) * (size_t)G));
    HIP_TRY(hipMalloc(&m->d_mid, sizeof(float) * kMlpParams * H));
    HIP_TRY(hipMalloc(&m->d_mid_obj, sizeof(double) * H));
    m->partial_groups = G;
  }
  MlpTrainArgs a{};
  a.net = m->net();
  a.sys = ctx->d_sysdev + sys;
  a.cfg = m->d_cfg;
  a.n_cfg = (uint32_t)m->n_cfg;
  a.first = first;
  a.batch = batch;
  a.seed = seed;
  a.partial = m->d_partial;
  a.obj = m->d_obj;
  a.objective = (int)objective;
  HIP_TRY(launch_mlp_train(a, G, m->stream));
  HIP_TRY(launch_mlp_reduce(m->d_partial, m->d_obj, G, m->d_mid, m->d_mid_obj, m->d_grad, obj_dev, adam, m->stream));
  return MOLE_OK;
}

mole_status mlp_train_prepare(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t batch) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!m) return fail(MOLE_ERR_INVALID_ARGUMENT, "null network");
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "training needs a Slater-Jastrow system");
  st = mlp_check_system(ctx->sys[sys]);
  if (st != MOLE_OK) return st;
  if (m->n_cfg == 0 || m->data_N != ctx->sys[sys].N)
    return fail(MOLE_ERR_SHAPE, "no training configurations of this system's electron count");
  if (batch == 0 || batch % 16 != 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 16");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->sysdev_dirty) {
    st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MOLE_OK;
}

}  // namespace

mole_status mole_mlp_gradient(mole_mlp *m, mole_ctx *ctx, uint32_t sys, uint32_t batch, uint64_t first,
                              uint64_t seed, uint32_t obj_kind, float *grad, double *objective) {
  if (obj_kind > MOLE_MLP_OBJ_LOG_ESJD) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown training objective %u", obj_kind);
  mole_status st = mlp_train_prepare(m, ctx, sys, batch);
  if (st != MOLE_OK) return st;
  if (!m->d_obj_steps) {
    HIP_TRY(hipMalloc(&m->d_obj_steps, sizeof(double) * 1024));
    m->obj_steps_cap = 1024;
  }
  st = mlp_gradient_enqueue(m, ctx, sys, batch, first, seed, m->d_obj_steps, nullptr, obj_kind);
  if (st != MOLE_OK) return st;
  if (grad) HIP_TRY(hipMemcpyAsync(grad, m->d_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
  double o = 0.0;
  HIP_TRY(hipMemcpyAsync(&o, m->d_obj_steps, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (objective) *objective = o;
  return ok();
}

mole_status mole_mlp_train(mole_mlp *m, mole_ctx *ctx, uint32_t sys, const mole_train_desc *d,
                           double *objective) {
  if (!d) return fail(MOLE_ERR_INVALID_ARGUMENT, "null training descriptor");
  mole_status st = mlp_train_prepare(m, ctx, sys, d->batch);
  if (st != MOLE_OK) return st;
  if (!(d->lr > 0.0f) || !(d->beta1 >= 0.0f && d->beta1 < 1.0f) || !(d->beta2 >= 0.0f && d->beta2 < 1.0f) ||
      !(d->eps > 0.0f))
    return fail(MOLE_ERR_INVALID_ARGUMENT, "Adam needs lr > 0, 0 <= beta < 1, eps > 0");
  if (d->objective > MOLE_MLP_OBJ_LOG_ESJD)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown training objective %u", d->objective);
  if (d->steps == 0) return ok();
  if ((int)d->steps > m->obj_steps_cap) {
    if (m->d_obj_steps) (void)hipFree(m->d_obj_steps);
    m->d_obj_steps = nullptr;
    HIP_TRY(hipMalloc(&m->d_obj_steps, sizeof(double) * std::max<uint32_t>(d->steps, 1024)));
    m->obj_steps_cap = (int)std::max<uint32_t>(d->steps, 1024);
  }
  for (uint32_t k = 0; k < d->steps; ++k) {
    const uint64_t first = (uint64_t)m->adam_step * d->batch;
    // the gradient reduction applies the Adam step (step counts from 1)
    const MlpAdamStep adam{m->d_w, m->d_m, m->d_v, 1.0f / (float)d->batch, d->lr, d->beta1, d->beta2, d->eps,
                           (int)m->adam_step + 1};
    st = mlp_gradient_enqueue(m, ctx, sys, d->batch, first, d->seed, m->d_obj_steps + k, &adam, d->objective);
    if (st != MOLE_OK) return st;
    ++m->adam_step;
    HIP_TRY(launch_mlp_pack(m->d_w, m->d_fwd1, m->d_fwd, m->d_bwd, m->d_bias, m->stream));
  }
  if (objective) {
    std::vector<double> o(d->steps);
    HIP_TRY(hipMemcpyAsync(o.data(), m->d_obj_steps, sizeof(double) * d->steps, hipMemcpyDeviceToHost,
                           m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    for (uint32_t k = 0; k < d->steps; ++k) objective[k] = o[k] / (double)d->batch;
  }
  HIP_TRY(hipStreamSynchronize(m->stream));
  return ok();
}

mole_status mole_mlp_eval(mole_mlp *m, const double *cfg, uint32_t N, const double *nuc_xyz, uint32_t n_nuc,
                          const int32_t *electron, uint32_t n, float *out) {
  if (!m || !cfg || !nuc_xyz || !electron || !out) return fail(MOLE_ERR_INVALID_ARGUMENT, "null argument");
  if (N == 0 || N > 64 || n_nuc == 0) return fail(MOLE_ERR_SHAPE, "need 1 <= N <= 64 and nuclei");
  for (uint32_t e = 0; e < n; ++e)
    if (electron[e] < 0 || (uint32_t)electron[e] >= N) return fail(MOLE_ERR_SHAPE, "electron index out of range");
  if (n == 0) return ok();
  HIP_TRY(hipSetDevice(m->device));
  double *dc = nullptr, *dn = nullptr;
  int *de = nullptr;
  float *dout = nullptr;
  auto cleanup = [&]() {
    if (dc) (void)hipFree(dc);
    if (dn) (void)hipFree(dn);
    if (de) (void)hipFree(de);
    if (dout) (void)hipFree(dout);
  };
  hipError_t e = hipMalloc(&dc, sizeof(double) * 3 * N * n);
  if (e == hipSuccess) e = hipMalloc(&dn, sizeof(double) * 3 * n_nuc);
  if (e == hipSuccess) e = hipMalloc(&de, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&dout, sizeof(float) * 4 * n);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, cfg, sizeof(double) * 3 * N * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dn, nuc_xyz, sizeof(double) * 3 * n_nuc, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(de, electron, sizeof(int) * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = launch_mlp_eval(m->net(), dc, (int)N, dn, (int)n_nuc, de, (int)n, dout, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * 4 * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  cleanup();
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_mlp_eval: %s", hipGetErrorString(e));
  return ok();
}

// ---- the optimisation loop on the device ------------------------------------

mole_status mole_optimize_with(mole_ctx *ctx, uint32_t iters, uint32_t steps, uint32_t block_size,
                               mole_optimizer *opt, double *log, double *params_out) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (ctx->sys.size() != 1) return fail(MOLE_ERR_UNSUPPORTED, "device-side optimisation covers one system");
  const bool sj = ctx->family == kFamSJ;
  if (block_size != ctx->run.block_size)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "block_size %u differs from the context's %u", block_size,
                ctx->run.block_size);
  if (steps < 2 * block_size)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "steps (%u) < 2 * block_size (%u)", steps, block_size);
  st = check_optimizer(opt);
  if (st != MOLE_OK) return st;
  if (opt->n_params != ctx->sys[0].params.size())
    return fail(MOLE_ERR_SHAPE, "optimizer has %u parameters, the system %zu", opt->n_params,
                ctx->sys[0].params.size());
  if (ctx->run.flags & (MOLE_FLAG_TRACE | MOLE_FLAG_SAMPLES))
    return fail(MOLE_ERR_UNSUPPORTED, "device-side optimisation records no traces or samples");
  if (iters == 0) return ok();
  SysState &s = ctx->sys[0];
  HIP_TRY(hipSetDevice(ctx->device));
  if (!ctx->d_params) HIP_TRY(hipMalloc(&ctx->d_params, sizeof(double) * MOLE_MAX_PARAMS));
  if ((size_t)iters > ctx->optlog_cap) {
    if (ctx->d_optlog) (void)hipFree(ctx->d_optlog);
    ctx->d_optlog = nullptr;
    HIP_TRY(hipMalloc(&ctx->d_optlog, sizeof(double) * 8 * iters));
    ctx->optlog_cap = iters;
  }
  if (!ctx->d_opt) HIP_TRY(hipMalloc(&ctx->d_opt, sizeof(mole_optimizer)));
  HIP_TRY(hipMemcpyAsync(ctx->d_params, s.params.data(), sizeof(double) * s.params.size(), hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->d_opt, opt, sizeof(mole_optimizer), hipMemcpyHostToDevice, ctx->stream));
  if (sj && ctx->sysdev_dirty) {
    // the sweep kernels read the parameters from the device system table,
    // which the step kernel updates in place: upload it once, before the loop
    st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  ctx->dev_params = !sj;
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  const uint32_t blocks = steps / block_size;
  uint32_t launches = 0;
  for (uint32_t it = 0; it < iters; ++it) {
    // reset statistics (as mole_reset_stats: the first sweep launch starts
    // the accumulators from zero; steps >= 2 blocks, so there is one),
    // Runner::run, reduce, step
    s.n_samples = s.n_blocks = s.n_moves = 0;
    ctx->block_fill = 0;
    ctx->trace_sweeps = ctx->trace_samples = 0;
    if (sj) {
      // as mole_reset_stats: zero accumulators and the launch state's block cursor
      HIP_TRY(launch_fill(s.d_acc, (size_t)s.n_cols * s.W, 0.0, ctx->stream));
      if (s.d_obs_acc) HIP_TRY(launch_fill(s.d_obs_acc, 2 * (size_t)MOLE_MAX_OBSERVABLES * s.W, 0.0, ctx->stream));
      st = sj_push_state(ctx);
      if (st != MOLE_OK) return st;
      launches += 2;
    } else {
      ctx->fresh_acc = true;
    }
    st = enqueue_sweeps(ctx, blocks * block_size, block_size, &launches);
    ctx->fresh_acc = false;
    if (st != MOLE_OK) {
      ctx->dev_params = false;
      return st;
    }
    if (sj) {
      HIP_TRY(launch_reduce_columns(s.d_acc, s.W, s.n_cols, ctx->d_red, kSJMaxFirst, kSJMaxCount, ctx->stream));
      HIP_TRY(launch_sj_vmc_step(ctx->d_red, (double)s.n_samples, (int)s.params.size(), ctx->d_params, ctx->d_opt,
                                 ctx->d_optlog + 8 * (size_t)it, ctx->d_sysdev, s.d_zeta0, s.d_tab + 4 * s.n_nuc,
                                 s.n_basis, ctx->stream));
    } else {
      HIP_TRY(launch_reduce_columns(s.d_acc, s.W, s.n_cols, ctx->d_red, -1, 0, ctx->stream));
      HIP_TRY(launch_vmc_step(ctx->d_red, (double)s.n_samples, ctx->d_params, ctx->d_opt,
                              ctx->d_optlog + 8 * (size_t)it, ctx->stream));
    }
    launches += 2;
  }
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  ctx->dev_params = false;
  ctx->last_launches = launches;
  ctx->timed = true;
  std::vector<double> p(s.params.size()), lg(8 * (size_t)iters);
  HIP_TRY(hipMemcpyAsync(p.data(), ctx->d_params, sizeof(double) * p.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(lg.data(), ctx->d_optlog, sizeof(double) * lg.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(opt, ctx->d_opt, sizeof(mole_optimizer), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  s.params = p;
  if (sj) {
    ctx->sysdev_dirty = true;  // the host copy of the system table follows
  } else {
    s.sp.zeta = p[0];
    s.sp.b = p[1];
    s.sp.c = p[2];
  }
  s.cols_valid = false;
  if (params_out) std::memcpy(params_out, p.data(), sizeof(double) * p.size());
  if (log) std::memcpy(log, lg.data(), sizeof(double) * lg.size());
  for (uint32_t it = 0; it < iters; ++it)
    if (lg[8 * (size_t)it + 5] != 0.0) return fail(MOLE_ERR_LINALG, "SR matrix singular at iteration %u", it);
  return ok();
}

mole_status mole_optimize(mole_ctx *ctx, uint32_t iters, uint32_t steps, uint32_t block_size, uint32_t optimizer,
                          double step, uint32_t sr_mode, double sr_eps, double *log, double *params_out) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (optimizer > MOLE_OPT_SR || sr_mode > MOLE_SR_REFERENCE)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown optimizer or SR mode");
  mole_optimizer opt;
  st = mole_optimizer_init(&opt, optimizer, (uint32_t)ctx->sys[0].params.size(), step, 0.0, 0, sr_mode, sr_eps);
  if (st != MOLE_OK) return st;
  return mole_optimize_with(ctx, iters, steps, block_size, &opt, log, params_out);
}

// ---- diffusion Monte Carlo --------------------------------------------------

namespace {

template <typename P>
hipError_t dmalloc(P **p, size_t n) {
  return hipMalloc((void **)p, sizeof(P) * (n ? n : 1));
}

mole_status dmc_alloc(mole_ctx *ctx, SysState &s) {
  SysState::Dmc &d = s.dmc;
  if (d.dmc) return MOLE_OK;
  const size_t W = (size_t)s.W;
  const size_t nB = (size_t)s.n_up * s.n_up + (size_t)s.n_down * s.n_down;
  HIP_TRY(dmalloc(&d.dmc, 4 * W));
  HIP_TRY(dmalloc(&d.st, 1));
  HIP_TRY(dmalloc(&d.part, 4 * dmc_parts(s.W)));
  HIP_TRY(dmalloc(&d.cnt_a, 3 * W));
  HIP_TRY(dmalloc(&d.cnt_b, W));
  HIP_TRY(dmalloc(&d.bsum_a, dmc_scan_blocks(3 * W)));
  HIP_TRY(dmalloc(&d.bsum_b, dmc_scan_blocks(W)));
  HIP_TRY(dmalloc(&d.list, 3 * W));
  HIP_TRY(dmalloc(&d.removed, 2 * W));
  HIP_TRY(dmalloc(&d.src, W));
  HIP_TRY(dmalloc(&d.wnew, W));
  HIP_TRY(dmalloc(&d.pos_s, 3 * (size_t)s.N * W));
  HIP_TRY(dmalloc(&d.el_s, W));
  HIP_TRY(hipMalloc(&d.B_s, (nB ? nB : 1) * W * (size_t)sj_tsize(ctx)));
  return MOLE_OK;
}

DmcArgs dmc_args(mole_ctx *ctx, uint32_t sys, uint32_t block_size, uint32_t num_eq_blocks) {
  SysState &s = ctx->sys[sys];
  SysState::Dmc &d = s.dmc;
  DmcArgs a{};
  a.st = d.st;
  a.dmc = d.dmc;
  a.W = s.W;
  a.brancher = (int)ctx->dmc_brancher;
  a.seed = ctx->run.seed;
  a.sys = sys;
  a.block_size = block_size;
  a.num_eq_blocks = num_eq_blocks;
  a.cap_steps = (uint32_t)d.cap_steps;
  a.step_energy = d.step_energy;
  a.block_buf = d.block_buf;
  a.energies = d.energies;
  a.vars = d.vars;
  a.part = d.part;
  a.cnt_a = d.cnt_a;
  a.cnt_b = d.cnt_b;
  a.bsum_a = d.bsum_a;
  a.bsum_b = d.bsum_b;
  a.list = d.list;
  a.removed = d.removed;
  a.src = d.src;
  a.wnew = d.wnew;
  a.pos = s.d_pos;
  a.pos_rows = 3 * (int)s.N;
  a.B = s.d_B;
  a.b_rows = (int)(s.n_up * s.n_up + s.n_down * s.n_down);
  a.b_elem = sj_tsize(ctx);
  a.pos_s = d.pos_s;
  a.B_s = d.B_s;
  a.el_s = d.el_s;
  return a;
}

// one DMC step of every system: the sweep launch, then each system's
// reductions, branching, gather and bookkeeping
mole_status dmc_enqueue_step(mole_ctx *ctx, double time_step, uint32_t block_size, uint32_t num_eq_blocks) {
  SJArgs a = sj_args(ctx);
  a.n_sweeps = 1;
  a.sample_start = 0;
  a.dmc = kSJDmcStep;
  a.dmc_tau = time_step;
  a.ext = 1;
  HIP_TRY(sj_launch_groups(ctx, a));
  for (uint32_t s = 0; s < ctx->sys.size(); ++s) {
    DmcArgs d = dmc_args(ctx, s, block_size, num_eq_blocks);
    if (s == 0) {  // the step's sweep, counted once (no samples)
      d.sweeps_done = &ctx->d_state->sweeps_done;
      d.trace_sweeps = &ctx->d_state->trace_sweeps;
    }
    HIP_TRY(launch_dmc_after_step(d, ctx->stream));
  }
  return MOLE_OK;
}

}  // namespace

mole_status mole_dmc_init(mole_ctx *ctx, const double *reference_energy, uint32_t brancher) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (!reference_energy) return fail(MOLE_ERR_INVALID_ARGUMENT, "null reference energies");
  if (ctx->family != kFamSJ || ctx->run.proposal != MOLE_PROPOSAL_DIFFUSE)
    return fail(MOLE_ERR_UNSUPPORTED, "DMC runs Slater-Jastrow systems with MOLE_PROPOSAL_DIFFUSE");
  if (brancher > MOLE_BRANCH_SIMPLE) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown brancher %u", brancher);
  HIP_TRY(hipSetDevice(ctx->device));
  for (SysState &s : ctx->sys) {
    st = dmc_alloc(ctx, s);
    if (st != MOLE_OK) return st;
  }
  // the move streams start over (as mole_set_stream(epoch, 0))
  for (SysState &s : ctx->sys) {
    s.draw_base = 0;
    s.move_base = 0;
    s.cols_valid = false;
  }
  ctx->sweeps_done = 0;
  ctx->sysdev_dirty = true;
  st = sj_upload_sysdev(ctx);
  if (st != MOLE_OK) return st;
  st = sj_push_state(ctx);
  if (st != MOLE_OK) return st;
  for (uint32_t k = 0; k < ctx->sys.size(); ++k) {
    SysState &s = ctx->sys[k];
    HIP_TRY(launch_dmc_init_configs(s.d_pos, (int)s.N, s.W, ctx->run.seed, k, ctx->stream));
    DmcState v{};
    v.eref = reference_energy[k];
    HIP_TRY(launch_dmc_state_set(s.dmc.st, v, ctx->stream));
  }
  // inverses of the configuration, local energies, unit weights
  SJArgs a = sj_args(ctx);
  a.n_sweeps = 0;
  a.refresh = 1;
  a.dmc = kSJDmcInit;
  a.ext = 1;
  HIP_TRY(sj_launch_groups(ctx, a));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->dmc_brancher = brancher;
  ctx->dmc_ready = true;
  return ok();
}

mole_status mole_dmc_diffuse(mole_ctx *ctx, double time_step, uint32_t num_iterations, uint32_t block_size,
                             uint32_t num_eq_blocks) {
  mole_status st = check_ctx(ctx, 0);
  if (st != MOLE_OK) return st;
  if (!ctx->dmc_ready) return fail(MOLE_ERR_INVALID_ARGUMENT, "call mole_dmc_init first");
  if (block_size == 0) return fail(MOLE_ERR_INVALID_ARGUMENT, "block_size must be >= 1");
  if (!std::isfinite(time_step)) return fail(MOLE_ERR_INVALID_ARGUMENT, "time_step must be finite");
  HIP_TRY(hipSetDevice(ctx->device));
  const uint32_t blocks = num_iterations / block_size;
  const size_t steps = (size_t)blocks * block_size;
  for (SysState &s : ctx->sys) {
    SysState::Dmc &d = s.dmc;
    if (steps > d.cap_steps) {
      if (d.step_energy) (void)hipFree(d.step_energy);
      d.step_energy = nullptr;
      HIP_TRY(dmalloc(&d.step_energy, steps));
      d.cap_steps = steps;
    }
    if (block_size > d.cap_block) {
      if (d.block_buf) (void)hipFree(d.block_buf);
      d.block_buf = nullptr;
      HIP_TRY(dmalloc(&d.block_buf, block_size));
      d.cap_block = block_size;
    }
    if (blocks > d.cap_blocks || !d.energies) {
      if (d.energies) (void)hipFree(d.energies);
      if (d.vars) (void)hipFree(d.vars);
      d.energies = d.vars = nullptr;
      HIP_TRY(dmalloc(&d.energies, blocks));
      HIP_TRY(dmalloc(&d.vars, blocks));
      d.cap_blocks = blocks;
    }
    HIP_TRY(launch_dmc_call_start(d.st, ctx->stream));
    d.steps_out = steps;
  }
  if (ctx->sysdev_dirty) {
    st = sj_upload_sysdev(ctx);
    if (st != MOLE_OK) return st;
  }
  // one step (sweep launch, then each system's reductions, branching and
  // bookkeeping) captured once and replayed: every count the kernels need is
  // in device memory
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (steps > 1) {
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    st = dmc_enqueue_step(ctx, time_step, block_size, num_eq_blocks);
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != MOLE_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return st != MOLE_OK ? st : fail(MOLE_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(MOLE_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
  }
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  // the steps start from the gathered copies of the walkers and leave the
  // last branching's choice there; the walker state is brought up to date
  // at the end
  for (uint32_t k = 0; k < ctx->sys.size(); ++k)
    HIP_TRY(launch_dmc_sync_state(dmc_args(ctx, k, block_size, num_eq_blocks), false, ctx->stream));
  hipError_t le = hipSuccess;
  for (size_t it = 0; it < steps && le == hipSuccess; ++it) {
    if (exec) {
      le = hipGraphLaunch(exec, ctx->stream);
    } else {
      st = dmc_enqueue_step(ctx, time_step, block_size, num_eq_blocks);
      if (st != MOLE_OK) return st;
    }
    account(ctx, 1, 0);
  }
  for (uint32_t k = 0; k < ctx->sys.size() && le == hipSuccess; ++k)
    le = launch_dmc_sync_state(dmc_args(ctx, k, block_size, num_eq_blocks), true, ctx->stream);
  if (le == hipSuccess) le = hipEventRecord(ctx->ev1, ctx->stream);
  if (le == hipSuccess) le = hipStreamSynchronize(ctx->stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (le != hipSuccess) return fail(MOLE_ERR_DEVICE, "DMC steps: %s", hipGetErrorString(le));
  ctx->last_launches = (uint32_t)steps;
  ctx->timed = true;
  {
    // the steps do not carry the log-determinant state of the walkers they
    // branch: take it from the walkers' configurations (inverses kept)
    SJArgs a = sj_args(ctx);
    a.n_sweeps = 0;
    a.refresh = 2;
    HIP_TRY(sj_launch_groups(ctx, a));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (uint32_t k = 0; k < ctx->sys.size(); ++k) {
    DmcState v;
    HIP_TRY(hipMemcpy(&v, ctx->sys[k].dmc.st, sizeof v, hipMemcpyDeviceToHost));
    if (v.error)
      return fail(MOLE_ERR_INVALID_ARGUMENT, "system %u: every walker's integer weight is 0 (SRBrancher)", k);
  }
  return ok();
}

mole_status mole_dmc_read(mole_ctx *ctx, uint32_t sys, double *energies, double *errors, size_t cap,
                          size_t *n_blocks, double *reference_energy) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!ctx->dmc_ready) return fail(MOLE_ERR_INVALID_ARGUMENT, "call mole_dmc_init first");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const SysState::Dmc &d = ctx->sys[sys].dmc;
  DmcState v;
  HIP_TRY(hipMemcpy(&v, d.st, sizeof v, hipMemcpyDeviceToHost));
  const size_t n = d.energies ? v.n_out : 0;
  if (n_blocks) *n_blocks = n;
  if (reference_energy) *reference_energy = v.eref;
  if ((energies || errors) && cap < n) return fail(MOLE_ERR_SHAPE, "capacity %zu < %zu blocks", cap, n);
  if (n == 0) return ok();
  std::vector<double> e(n), var(n);
  HIP_TRY(hipMemcpy(e.data(), d.energies, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(var.data(), d.vars, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (energies) std::memcpy(energies, e.data(), sizeof(double) * n);
  if (errors)
    for (size_t i = 0; i < n; ++i) errors[i] = std::sqrt(var[i] / (double)(i + 1));  // dmc.rs:147-150
  return ok();
}

mole_status mole_dmc_read_steps(mole_ctx *ctx, uint32_t sys, double *step_energy, size_t cap, size_t *n) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!ctx->dmc_ready) return fail(MOLE_ERR_INVALID_ARGUMENT, "call mole_dmc_init first");
  const SysState::Dmc &d = ctx->sys[sys].dmc;
  if (n) *n = d.steps_out;
  if (!step_energy || d.steps_out == 0) return ok();
  if (cap < d.steps_out) return fail(MOLE_ERR_SHAPE, "capacity %zu < %zu steps", cap, d.steps_out);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(step_energy, d.step_energy, sizeof(double) * d.steps_out, hipMemcpyDeviceToHost));
  return ok();
}

mole_status mole_dmc_read_walkers(mole_ctx *ctx, uint32_t sys, double *weights, double *e_local, size_t n_walkers) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (!ctx->dmc_ready) return fail(MOLE_ERR_INVALID_ARGUMENT, "call mole_dmc_init first");
  const SysState &s = ctx->sys[sys];
  if (n_walkers != (size_t)s.W) return fail(MOLE_ERR_SHAPE, "system %u has %d walkers", sys, s.W);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (weights) HIP_TRY(hipMemcpy(weights, s.dmc.dmc, sizeof(double) * s.W, hipMemcpyDeviceToHost));
  if (e_local) HIP_TRY(hipMemcpy(e_local, s.dmc.dmc + s.W, sizeof(double) * s.W, hipMemcpyDeviceToHost));
  return ok();
}

// ---- local operators -----------------------------------------------------------

namespace {

mole_status check_operator(const mole_operator &op) {
  if (op.n_terms < 1 || op.n_terms > MOLE_OPERATOR_MAX_TERMS)
    return fail(MOLE_ERR_INVALID_ARGUMENT, "an operator has 1..%d terms", MOLE_OPERATOR_MAX_TERMS);
  for (uint32_t k = 0; k < op.n_terms; ++k) {
    const mole_term &t = op.term[k];
    if (t.kind > MOLE_TERM_DIPOLE) return fail(MOLE_ERR_INVALID_ARGUMENT, "unknown term kind %u", t.kind);
    if (t.kind == MOLE_TERM_RADIAL_POWER) {
      const double p = t.param[0];
      if (!(p == std::floor(p)) || p == 0.0 || p < -4.0 || p > 4.0)
        return fail(MOLE_ERR_INVALID_ARGUMENT, "RADIAL_POWER needs an integer power in -4..4, not 0");
    }
  }
  return MOLE_OK;
}

mole_status push_ops(mole_ctx *ctx, SysState &s) {
  HIP_TRY(hipSetDevice(ctx->device));
  if (!s.d_ops) HIP_TRY(hipMalloc(&s.d_ops, sizeof(mole_operator) * (1 + MOLE_MAX_OBSERVABLES)));
  if (!s.d_obs_acc) {
    HIP_TRY(hipMalloc(&s.d_obs_acc, sizeof(double) * 2 * MOLE_MAX_OBSERVABLES * (size_t)s.W));
    HIP_TRY(launch_fill(s.d_obs_acc, 2 * (size_t)MOLE_MAX_OBSERVABLES * s.W, 0.0, ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(s.d_ops, s.ops.data(), sizeof(mole_operator) * s.ops.size(), hipMemcpyHostToDevice));
  ctx->sysdev_dirty = true;
  return MOLE_OK;
}

}  // namespace

mole_status mole_set_hamiltonian(mole_ctx *ctx, uint32_t sys, const mole_operator *h) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "local operators run on Slater-Jastrow systems");
  SysState &s = ctx->sys[sys];
  if (h) {
    st = check_operator(*h);
    if (st != MOLE_OK) return st;
    s.ops[0] = *h;
  }
  s.custom_h = h != nullptr;
  return push_ops(ctx, s);
}

mole_status mole_set_observables(mole_ctx *ctx, uint32_t sys, const mole_operator *obs, uint32_t n) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  if (ctx->family != kFamSJ) return fail(MOLE_ERR_UNSUPPORTED, "local operators run on Slater-Jastrow systems");
  if (n > MOLE_MAX_OBSERVABLES) return fail(MOLE_ERR_INVALID_ARGUMENT, "at most %d observables", MOLE_MAX_OBSERVABLES);
  if (n && !obs) return fail(MOLE_ERR_INVALID_ARGUMENT, "null observables");
  for (uint32_t k = 0; k < n; ++k) {
    st = check_operator(obs[k]);
    if (st != MOLE_OK) return st;
  }
  SysState &s = ctx->sys[sys];
  s.ops.resize(1 + n);
  for (uint32_t k = 0; k < n; ++k) s.ops[1 + k] = obs[k];
  s.n_obs = (int)n;
  st = push_ops(ctx, s);
  if (st != MOLE_OK) return st;
  HIP_TRY(launch_fill(s.d_obs_acc, 2 * (size_t)MOLE_MAX_OBSERVABLES * s.W, 0.0, ctx->stream));
  return ok();
}

mole_status mole_read_observables(mole_ctx *ctx, uint32_t sys, double *sum, double *sum2, uint32_t n,
                                  uint64_t *n_samples) {
  mole_status st = check_ctx(ctx, sys);
  if (st != MOLE_OK) return st;
  SysState &s = ctx->sys[sys];
  if (n != (uint32_t)s.n_obs) return fail(MOLE_ERR_SHAPE, "system %u has %d observables", sys, s.n_obs);
  if (n_samples) *n_samples = s.n_samples;
  if (n == 0) return ok();
  HIP_TRY(hipSetDevice(ctx->device));
  double *d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * n));
  hipError_t e = launch_reduce_columns(s.d_obs_acc, s.W, (int)(2 * n), d, -1, 0, ctx->stream);
  std::vector<double> h(2 * n);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(MOLE_ERR_DEVICE, "mole_read_observables: %s", hipGetErrorString(e));
  for (uint32_t k = 0; k < n; ++k) {
    if (sum) sum[k] = h[2 * k];
    if (sum2) sum2[k] = h[2 * k + 1];
  }
  return ok();
}
