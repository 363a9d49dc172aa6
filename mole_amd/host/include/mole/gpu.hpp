// Device-backed implementations of the reference traits over the C ABI of
// libmole_amd.so (include/mole_gpu.h):
//   gpu::Context          owns a mole_ctx (walkers of one or more systems)
//   GpuWaveFunction       Function + Differentiate + WaveFunction + Optimize
//                         of the two-electron kinds, evaluated on the device
//   GpuSampler<WF>        MonteCarloSampler + FusedSampler: walkers on the
//                         device, sweeps fused into launches
//   SteepestDescent, StochasticReconfiguration
//                         Optimizer over the device-reduced sums
// The call each one replaces is named at the class.
#pragma once

#include <functional>

#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "mole_gpu.h"
#include "traits.hpp"

namespace mole {

// A sampler whose sweeps run fused on the device: Runner::run and
// VmcRunner::run_optimization take this path (the results equal the
// per-move path's; only the launches differ).
struct FusedSampler {
  virtual ~FusedSampler() = default;
  // steps sweeps in blocks of block_size, the first block discarded
  // (montecarlo.rs:24-46); statistics restart at the call
  virtual Status run_fused(size_t steps, size_t block_size) = 0;
  // the same, calling after_block() once every sampled block has finished
  // with data() holding the blocks so far (the cadence at which Runner::run
  // prints Log::log's output, montecarlo.rs:31-42); false if the sampler
  // does not stream blocks
  virtual bool streams_blocks() const { return false; }
  virtual Status run_fused_streaming(size_t steps, size_t block_size, const std::function<void()> &after_block) {
    (void)after_block;
    return run_fused(steps, block_size);
  }
  // energy mean, blocking error and the SR sums of the last run
  virtual Result<mole_stats> stats() const = 0;
  virtual Result<mole_sr_sums> sr_sums() const = 0;
};

namespace gpu {

class Context {
 public:
  Context() = default;
  Context(const std::vector<mole_system_desc> &sys, const mole_run_desc &run) {
    mole_ctx *c = nullptr;
    const Status st = status(mole_ctx_create(sys.data(), (uint32_t)sys.size(), &run, &c));
    if (st.is_err()) {
      err_ = std::make_unique<Error>(st.error());
      return;
    }
    ctx_.reset(c);
  }
  bool ok() const { return (bool)ctx_; }
  const Error *error() const { return err_.get(); }
  mole_ctx *get() const { return ctx_.get(); }

 private:
  struct Del {
    void operator()(mole_ctx *c) const { mole_ctx_destroy(c); }
  };
  std::unique_ptr<mole_ctx, Del> ctx_;
  std::unique_ptr<Error> err_;
};

inline mole_run_desc run_desc(uint64_t seed, double step, uint32_t block_size, uint32_t flags = 0,
                              uint32_t record_sweeps = 0) {
  mole_run_desc r{};
  r.seed = seed;
  r.proposal = MOLE_PROPOSAL_BOX;
  r.step = step;
  r.block_size = block_size;
  r.o_convention = MOLE_O_LOGDERIV;
  r.flags = flags;
  r.record_sweeps = record_sweeps;
  r.precision = MOLE_PREC_FP64;
  return r;
}

}  // namespace gpu

// Two-electron wavefunctions of the examples (He singlet with Pade
// Jastrow, examples/helium_atom_singlet.rs:34-118; H2 LCAO with Pade
// Jastrow) with parameters (zeta, b, c), evaluated by mole_eval.  Copies
// share nothing: a copy evaluates on its own one-walker context (Clone in
// the reference).
class GpuWaveFunction : public Function, public Differentiate, public WaveFunction, public Optimize {
 public:
  GpuWaveFunction(uint32_t kind, std::vector<double> nuclei_xyz, std::vector<int32_t> charges, Array1 params,
                  uint32_t hamiltonian = MOLE_H_ELECTRONIC)
      : kind_(kind), hamiltonian_(hamiltonian), nuc_(std::move(nuclei_xyz)), z_(std::move(charges)),
        params_(std::move(params)) {}
  GpuWaveFunction(const GpuWaveFunction &o)
      : kind_(o.kind_), hamiltonian_(o.hamiltonian_), nuc_(o.nuc_), z_(o.z_), params_(o.params_) {}
  GpuWaveFunction &operator=(const GpuWaveFunction &o) {
    kind_ = o.kind_;
    hamiltonian_ = o.hamiltonian_;
    nuc_ = o.nuc_;
    z_ = o.z_;
    params_ = o.params_;
    eval_.reset();
    return *this;
  }

  static GpuWaveFunction helium(double zeta, double b, double c) {
    return GpuWaveFunction(MOLE_WF_HELIUM_PADE, {0.0, 0.0, 0.0}, {2}, {zeta, b, c});
  }
  static GpuWaveFunction hydrogen_molecule(double zeta, double b, double c, double half_bond = 0.7) {
    return GpuWaveFunction(MOLE_WF_LCAO2_PADE, {0.0, 0.0, -half_bond, 0.0, 0.0, half_bond}, {1, 1}, {zeta, b, c});
  }

  // mole_system_desc of this wavefunction for W walkers (pointers into *this)
  mole_system_desc desc(uint32_t walkers) const {
    mole_system_desc d{};
    d.wf_kind = kind_;
    d.hamiltonian = hamiltonian_;
    d.n_up = 1;
    d.n_down = 1;
    d.n_nuclei = (uint32_t)z_.size();
    d.nuclei_xyz = nuc_.data();
    d.nuclei_charge = z_.data();
    d.n_params = (uint32_t)params_.size();
    d.params = params_.data();
    d.n_walkers = walkers;
    return d;
  }

  // Function::value
  Result<double> value(const Array2 &cfg) const override {
    double v = 0.0;
    const Status st = eval(cfg, &v, nullptr, nullptr, nullptr, nullptr);
    if (st.is_err()) return st.error();
    return v;
  }
  // Differentiate::gradient (d psi, rows = electrons)
  Result<Array2> gradient(const Array2 &cfg) const override {
    Array2 g(2, 3);
    const Status st = eval(cfg, nullptr, g.data.data(), nullptr, nullptr, nullptr);
    if (st.is_err()) return st.error();
    return g;
  }
  Result<double> laplacian(const Array2 &cfg) const override {
    double l = 0.0;
    const Status st = eval(cfg, nullptr, nullptr, &l, nullptr, nullptr);
    if (st.is_err()) return st.error();
    return l;
  }
  // ElectronicHamiltonian::act_on(psi, cfg) / psi (operator.rs:177-183)
  Result<double> local_energy(const Array2 &cfg) const {
    double e = 0.0;
    const Status st = eval(cfg, nullptr, nullptr, nullptr, &e, nullptr);
    if (st.is_err()) return st.error();
    return e;
  }
  size_t num_electrons() const override { return 2; }
  // Optimize (src/optimize/src/traits.rs:8-16): d psi / dp
  Result<Array1> parameter_gradient(const Array2 &cfg) const override {
    Array1 g(params_.size());
    const Status st = eval(cfg, nullptr, nullptr, nullptr, nullptr, g.data());
    if (st.is_err()) return st.error();
    return g;
  }
  void update_parameters(const Array1 &deltap) override {
    for (size_t k = 0; k < params_.size() && k < deltap.size(); ++k) params_[k] += deltap[k];
    eval_.reset();
  }
  const Array1 &parameters() const override { return params_; }
  size_t num_parameters() const override { return params_.size(); }

 private:
  Status eval(const Array2 &cfg, double *v, double *g, double *l, double *e, double *pg) const {
    if (cfg.rows != 2 || cfg.cols != 3) return Error{ErrorKind::ShapeError, "expected a 2 x 3 configuration"};
    if (!eval_) {
      eval_ = std::make_shared<gpu::Context>(std::vector<mole_system_desc>{desc(1)}, gpu::run_desc(0, 0.5, 1));
      if (!eval_->ok()) {
        const Error err = *eval_->error();
        eval_.reset();
        return err;
      }
    }
    return status(mole_eval(eval_->get(), 0, 1, cfg.data.data(), v, g, l, e, pg));
  }

  uint32_t kind_, hamiltonian_;
  std::vector<double> nuc_;
  std::vector<int32_t> z_;
  Array1 params_;
  mutable std::shared_ptr<gpu::Context> eval_;
};

// Sampler::new + Sampler::move_state/sample (src/montecarlo/src/samplers.rs:
// 40-117) for W walkers of one system on the device: every walker is a
// chain, a sweep moves each electron once with MetropolisBox
// (metrop.rs:60-96) on its own Philox stream.  move_state()/sample() follow
// the reference per sweep (one launch each); run_fused() runs a whole
// Runner::run in launches of many sweeps.  data() holds the reduced
// observables: "Energy" [mean], "Energy error" [blocking error],
// "SR sums" [1 x (2 + 2p + p^2) matrix: n, sum_e, sum_o, sum_oe, sum_oo].
template <typename WF>
class GpuSampler : public MonteCarloSampler<WF>, public FusedSampler {
 public:
  // record_sweeps > 0 keeps the local-energy trace of that many samples
  // (MOLE_FLAG_TRACE), so data() also holds "Energy block means" (one
  // scalar per finished block, mole_read_blocks) and runs stream blocks
  GpuSampler(WF wf, uint32_t walkers, double box_side, uint64_t seed, uint32_t block_size = 1,
             uint32_t record_sweeps = 0)
      : wf_(std::move(wf)), walkers_(walkers), side_(box_side), seed_(seed), block_(block_size),
        record_(record_sweeps) {
    open();
  }

  bool ok() const { return ctx_ && ctx_->ok(); }
  const Error *error() const { return ctx_ ? ctx_->error() : nullptr; }
  mole_ctx *context() const { return ctx_->get(); }

  Status move_state() override {
    Status st = flush();
    if (st.is_err()) return st;
    pending_ = true;
    return Ok();
  }
  Status sample() override {
    if (!pending_) return Error{ErrorKind::DataAccessError, "sample() without a preceding move_state()"};
    pending_ = false;
    Status st = push_params();
    if (st.is_err()) return st;
    st = status(mole_sweep(ctx_->get(), 1, MOLE_SWEEP_SAMPLE));
    if (st.is_err()) return st;
    return refresh_data();
  }
  const DataMap &data() const override { return data_; }
  size_t num_observables() const override { return data_.size(); }
  // the reference's accumulator: 1/N per accepted move (samplers.rs:113),
  // summed over the walkers, since the last statistics reset
  double acceptance() const override {
    mole_stats s{};
    if (mole_read_stats(ctx_->get(), 0, &s) != MOLE_OK) return 0.0;
    return (double)s.n_accepted / (double)wf_.num_electrons();
  }
  std::vector<std::string> observable_names() const override {
    std::vector<std::string> n;
    for (const auto &kv : data_) n.push_back(kv.first);
    return n;
  }
  MonteCarloResult<WF> consume_result() override {
    (void)flush();
    return MonteCarloResult<WF>{wf_, acceptance(), data_};
  }
  const WF &wave_function() const override { return wf_; }
  WF &wave_function_mut() override { return wf_; }
  // the Philox key replaces the 32-byte StdRng seed (metrop.rs:43, 87-95)
  // the walkers keep their configurations; their streams restart
  void reseed_rng(const Seed &seed) override {
    uint64_t k = 0;
    std::memcpy(&k, seed.data(), sizeof k);
    seed_ = k;
    open_keep_state(true);
  }
  Seed generate_seed() override {
    Seed s{};
    const uint64_t k = seed_ * 6364136223846793005ull + 1442695040888963407ull;
    std::memcpy(s.data(), &k, sizeof k);
    return s;
  }

  Status run_fused(size_t steps, size_t block_size) override {
    if (steps < 2 * block_size) throw Panic("assertion failed: steps >= 2 * block_size");
    if (block_size != block_) {
      block_ = (uint32_t)block_size;
      open_keep_state(false);
    }
    Status st = flush();
    if (st.is_err()) return st;
    st = push_params();
    if (st.is_err()) return st;
    st = status(mole_reset_stats(ctx_->get()));
    if (st.is_err()) return st;
    st = status(mole_run(ctx_->get(), (uint32_t)steps, (uint32_t)block_size));
    if (st.is_err()) return st;
    return refresh_data();
  }
  // the "Energy" operator (mole_set_hamiltonian): the operators! "Energy"
  // entry of a run with a user LocalOperator made of device terms
  Status set_hamiltonian(const mole_operator &h) {
    h_ = h;
    has_h_ = true;
    return status(mole_set_hamiltonian(ctx_->get(), 0, &h_));
  }
  bool streams_blocks() const override { return record_ > 0; }
  Status run_fused_streaming(size_t steps, size_t block_size, const std::function<void()> &after_block) override {
    if (!record_) return run_fused(steps, block_size);
    if (steps < 2 * block_size) throw Panic("assertion failed: steps >= 2 * block_size");
    if (steps - block_size > record_)
      return Error{ErrorKind::ShapeError, "more samples than the sampler records"};
    if (block_size != block_) {
      block_ = (uint32_t)block_size;
      open_keep_state(false);
    }
    Status st = flush();
    if (st.is_err()) return st;
    st = push_params();
    if (st.is_err()) return st;
    st = status(mole_reset_stats(ctx_->get()));
    if (st.is_err()) return st;
    // the first block unsampled, then one launch per block
    st = status(mole_sweep(ctx_->get(), (uint32_t)block_size, 0));
    if (st.is_err()) return st;
    for (size_t b = 1; b < steps / block_size; ++b) {
      st = status(mole_sweep(ctx_->get(), (uint32_t)block_size, MOLE_SWEEP_SAMPLE));
      if (st.is_err()) return st;
      st = refresh_data();
      if (st.is_err()) return st;
      if (after_block) after_block();
    }
    return Ok();
  }
  // the per-block records since the last statistics reset
  Result<std::vector<mole_block_record>> block_records() const {
    size_t n = 0;
    Status st = status(mole_read_blocks(ctx_->get(), 0, 0, nullptr, 0, &n));
    if (st.is_err()) return st.error();
    std::vector<mole_block_record> r(n);
    if (n) {
      st = status(mole_read_blocks(ctx_->get(), 0, 0, r.data(), n, &n));
      if (st.is_err()) return st.error();
    }
    return r;
  }
  Result<mole_stats> stats() const override {
    mole_stats s{};
    const Status st = status(mole_read_stats(ctx_->get(), 0, &s));
    if (st.is_err()) return st.error();
    return s;
  }
  Result<mole_sr_sums> sr_sums() const override {
    mole_sr_sums s{};
    const Status st = status(mole_read_sr_sums(ctx_->get(), 0, &s));
    if (st.is_err()) return st.error();
    return s;
  }

 private:
  void open() {
    desc_ = wf_.desc(walkers_);
    ctx_ = std::make_shared<gpu::Context>(std::vector<mole_system_desc>{desc_}, run());
    if (ctx_->ok()) (void)mole_init_walkers(ctx_->get());
  }
  // a new block size or seed needs a new context; the walkers carry over
  // (and, unless restart_stream, their stream position)
  void open_keep_state(bool restart_stream) {
    std::vector<double> cfg((size_t)walkers_ * 3 * wf_.num_electrons());
    (void)mole_get_configs(ctx_->get(), 0, cfg.data(), walkers_);
    uint32_t epoch = 0;
    uint64_t pos = 0;
    (void)mole_get_stream(ctx_->get(), &epoch, &pos);
    desc_ = wf_.desc(walkers_);
    ctx_ = std::make_shared<gpu::Context>(std::vector<mole_system_desc>{desc_}, run());
    if (ctx_->ok()) {
      (void)mole_set_configs(ctx_->get(), 0, cfg.data(), walkers_);
      if (!restart_stream) (void)mole_set_stream(ctx_->get(), epoch, pos);
      if (has_h_) (void)mole_set_hamiltonian(ctx_->get(), 0, &h_);
    }
  }
  mole_run_desc run() const {
    return gpu::run_desc(seed_, side_, block_, record_ ? MOLE_FLAG_TRACE : 0u, record_);
  }
  Status flush() {
    if (!pending_) return Ok();
    pending_ = false;
    Status st = push_params();
    if (st.is_err()) return st;
    return status(mole_sweep(ctx_->get(), 1, 0));
  }
  Status push_params() {
    const Array1 &p = wf_.parameters();
    return status(mole_set_params(ctx_->get(), 0, p.data(), (uint32_t)p.size()));
  }
  Status refresh_data() {
    mole_stats s{};
    Status st = status(mole_read_stats(ctx_->get(), 0, &s));
    if (st.is_err()) return st;
    mole_sr_sums r{};
    st = status(mole_read_sr_sums(ctx_->get(), 0, &r));
    if (st.is_err()) return st;
    data_.clear();
    if (s.n_samples) {
      const double mean = s.sum_e / (double)s.n_samples;
      data_["Energy"] = {OperatorValue::scalar(mean)};
      if (s.n_blocks > 1) {
        const double bm = s.sum_block_mean / (double)s.n_blocks;
        const double v = s.sum_block_mean2 / (double)s.n_blocks - bm * bm;
        data_["Energy error"] = {OperatorValue::scalar(std::sqrt(v > 0.0 ? v : 0.0) /
                                                       std::sqrt((double)(s.n_blocks - 1)))};
      }
    }
    data_["SR sums"] = {OperatorValue::matrix(pack_sums(r))};
    if (record_) {
      const Result<std::vector<mole_block_record>> recs = block_records();
      if (recs.is_err()) return recs.error();
      std::vector<OperatorValue> bm;
      for (const mole_block_record &b : recs.value()) bm.push_back(OperatorValue::scalar(b.mean));
      data_["Energy block means"] = bm;
    }
    return Ok();
  }

 public:
  static Matrix pack_sums(const mole_sr_sums &r) {
    const size_t p = r.n_params;
    Matrix m{1, 2 + 2 * p + p * p, {}};
    m.data.push_back((double)r.n);
    m.data.push_back(r.sum_e);
    for (size_t k = 0; k < p; ++k) m.data.push_back(r.sum_o[k]);
    for (size_t k = 0; k < p; ++k) m.data.push_back(r.sum_oe[k]);
    for (size_t k = 0; k < p * p; ++k) m.data.push_back(r.sum_oo[k]);
    return m;
  }
  static Result<mole_sr_sums> unpack_sums(const Matrix &m) {
    mole_sr_sums r{};
    if (m.rows != 1 || m.data.size() < 2) return Error{ErrorKind::ShapeError, "malformed SR sums"};
    size_t p = 0;
    while (2 + 2 * p + p * p < m.data.size()) ++p;
    if (2 + 2 * p + p * p != m.data.size() || p > MOLE_MAX_PARAMS)
      return Error{ErrorKind::ShapeError, "malformed SR sums"};
    r.n_params = (uint32_t)p;
    r.n = (uint64_t)m.data[0];
    r.sum_e = m.data[1];
    for (size_t k = 0; k < p; ++k) r.sum_o[k] = m.data[2 + k];
    for (size_t k = 0; k < p; ++k) r.sum_oe[k] = m.data[2 + p + k];
    for (size_t k = 0; k < p * p; ++k) r.sum_oo[k] = m.data[2 + 2 * p + k];
    return r;
  }

 private:
  WF wf_;
  uint32_t walkers_;
  double side_;
  uint64_t seed_;
  uint32_t block_;
  uint32_t record_ = 0;
  mole_operator h_{};
  bool has_h_ = false;
  mole_system_desc desc_{};
  std::shared_ptr<gpu::Context> ctx_;
  DataMap data_;
  bool pending_ = false;
};

// The optimizers of src/optimize/src/optimizers.rs on the "SR sums" the
// device reduced (mole_optimizer_update): SteepestDescent (:9-30),
// MomentumDescent (:32-61), NesterovMomentum (:63-95), OnlineLbfgs
// (:97-179) and StochasticReconfiguration (:181-253).  The state is a
// mole_optimizer (plain data), so copies are independent optimizers as the
// reference's Clone makes them; state() hands it to mole_optimize_with.
class SumsOptimizer : public Optimizer {
 public:
  Result<Array1> compute_parameter_update(const Array1 &pars, const AverageMap &averages,
                                          const DataMap &raw_data) override {
    (void)averages;
    const auto it = raw_data.find("SR sums");
    if (it == raw_data.end() || it->second.empty())
      return Error{ErrorKind::DataAccessError, "no \"SR sums\" observable"};
    const Result<Matrix> m = it->second.back().get_matrix();
    if (m.is_err()) return m.error();
    const Result<mole_sr_sums> r = GpuSampler<GpuWaveFunction>::unpack_sums(m.value());
    if (r.is_err()) return r.error();
    if (r.value().n_params != pars.size()) return Error{ErrorKind::ShapeError, "parameter count mismatch"};
    if (!ready_) {  // SD and SR learn the parameter count from the first call
      const Status st = status(mole_optimizer_init(&opt_, kind_, (uint32_t)pars.size(), step_, mu_, history_,
                                                   sr_mode_, eps_));
      if (st.is_err()) return st.error();
      ready_ = true;
    }
    Array1 dp(pars.size()), g(pars.size());
    const Status st = status(mole_optimizer_update(&opt_, &r.value(), pars.data(), dp.data(), g.data()));
    if (st.is_err()) return st.error();
    last_gradient_ = g;
    return dp;
  }
  const Array1 &last_gradient() const { return last_gradient_; }
  // the state (valid once ready(): after the first update, or from the
  // constructor for the optimizers that take nparm)
  mole_optimizer &state() { return opt_; }
  bool ready() const { return ready_; }

 protected:
  SumsOptimizer(uint32_t kind, double step, double mu, uint32_t history, uint32_t sr_mode, double sr_eps,
                size_t nparm)
      : kind_(kind), step_(step), mu_(mu), history_(history), sr_mode_(sr_mode), eps_(sr_eps) {
    if (nparm) {
      const Status st = status(mole_optimizer_init(&opt_, kind, (uint32_t)nparm, step, mu, history, sr_mode, sr_eps));
      if (st.is_err()) throw Panic(st.error().message);
      ready_ = true;
    }
  }

 private:
  uint32_t kind_;
  double step_, mu_;
  uint32_t history_, sr_mode_;
  double eps_;
  bool ready_ = false;
  mole_optimizer opt_{};
  Array1 last_gradient_;
};

struct SteepestDescent : SumsOptimizer {
  explicit SteepestDescent(double step) : SumsOptimizer(MOLE_OPT_SD, step, 0.0, 0, MOLE_SR_TEXTBOOK, 0.0, 0) {}
};

struct MomentumDescent : SumsOptimizer {
  MomentumDescent(double step, double momentum_parameter, size_t nparm)
      : SumsOptimizer(MOLE_OPT_MOMENTUM, step, momentum_parameter, 0, 0, 0.0, nparm) {}
};

struct NesterovMomentum : SumsOptimizer {
  NesterovMomentum(double step, double momentum_parameter, size_t nparm)
      : SumsOptimizer(MOLE_OPT_NESTEROV, step, momentum_parameter, 0, 0, 0.0, nparm) {}
};

struct OnlineLbfgs : SumsOptimizer {
  // history: pairs kept, 1..MOLE_LBFGS_MAX_HISTORY (each call pushes its
  // pair twice, as optimizers.rs:163,170 do)
  OnlineLbfgs(double step, size_t history, size_t nparm)
      : SumsOptimizer(MOLE_OPT_LBFGS, step, 0.0, (uint32_t)history, 0, 0.0, nparm) {}
};

struct StochasticReconfiguration : SumsOptimizer {
  // sr_mode MOLE_SR_REFERENCE reproduces optimizers.rs:219-231 as written
  // (scalar subtraction, diag *= 1.01); MOLE_SR_TEXTBOOK subtracts Obar Obar^T
  // and scales the diagonal by 1 + eps
  StochasticReconfiguration(double step, uint32_t sr_mode = MOLE_SR_TEXTBOOK, double eps = 1e-3)
      : SumsOptimizer(MOLE_OPT_SR, step, 0.0, 0, sr_mode, eps, 0) {}
};

}  // namespace mole
