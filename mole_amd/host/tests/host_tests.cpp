// Tests of the C++ host layer over libmole_amd.so.  Without arguments: the
// parts that need no GPU (OperatorValue arithmetic as the reference defines
// it, status -> Error mapping, Runner and VmcRunner bookkeeping on a host
// sampler).  With --gpu: a device VMC optimisation of H2 through GpuSampler,
// StochasticReconfiguration and VmcRunner (one JSON line for the caller to
// compare with the Python path), and GpuWaveFunction evaluations.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mole/errors.hpp"
#include "mole/gpu.hpp"
#include "mole/dmc.hpp"
#include "mole/operator_value.hpp"
#include "mole/slater_jastrow.hpp"
#include "mole/runner.hpp"
#include "mole/traits.hpp"

using namespace mole;

static int g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                                         \
    }                                                                   \
  } while (0)

template <typename F>
static bool panics(F f) {
  try {
    f();
  } catch (const Panic &) {
    return true;
  }
  return false;
}

static void test_operator_value() {
  using OV = OperatorValue;
  const OV s = OV::scalar(2.0), v = OV::vector({1.0, 4.0});
  CHECK((s + s).get_scalar().value() == 4.0);
  // Vector / Scalar = Scalar ./ Vector (traits.rs:150)
  const auto q = (v / s).get_vector().value();
  CHECK(q[0] == 2.0 && q[1] == 0.5);
  // Vector - Scalar = Scalar - Vector; Vector - Vector reversed (traits.rs:102-103)
  const auto d = (v - s).get_vector().value();
  CHECK(d[0] == 1.0 && d[1] == -2.0);
  const auto dd = (v - OV::vector({3.0, 3.0})).get_vector().value();
  CHECK(dd[0] == 2.0 && dd[1] == -1.0);
  CHECK((v * s).get_vector().value()[1] == 8.0);
  CHECK(s.get_vector().is_err() && s.get_vector().error().kind == ErrorKind::OperatorValueAccessError);
  CHECK(panics([&] { (void)(v / v); }));
  CHECK(panics([&] { (void)(v + OV::matrix(Matrix{1, 2, {1.0, 2.0}})); }));
  CHECK(v.map([](double x) { return x * x; }).get_vector().value()[1] == 16.0);
}

static void test_status_mapping() {
  // validation runs before any device call, so this needs no GPU
  GpuWaveFunction he = GpuWaveFunction::helium(1.6875, 0.5, 0.3);
  mole_system_desc d = he.desc(16);
  d.n_up = 2;  // the two-electron kinds need n_up = n_down = 1
  mole_ctx *c = nullptr;
  const mole_run_desc run = gpu::run_desc(0, 0.5, 10);
  const Status st = status(mole_ctx_create(&d, 1, &run, &c));
  CHECK(st.is_err() && st.error().kind == ErrorKind::ShapeError);
  // a non-positive box side is the reference's Range::new assert: Panic
  mole_system_desc ok = he.desc(16);
  const mole_run_desc bad = gpu::run_desc(0, 0.0, 10);
  CHECK(panics([&] { (void)status(mole_ctx_create(&ok, 1, &bad, &c)); }));
  CHECK(error_from_status(MOLE_ERR_DEVICE).kind == ErrorKind::FuncError);
  CHECK(error_from_status(MOLE_ERR_LINALG).kind == ErrorKind::LinalgError);
}

// A host sampler of known energies: sample k of a run gives E = k, and the
// wavefunction's parameters are a plain vector.
struct HostWave : Optimize {
  Array1 p{1.0};
  Result<Array1> parameter_gradient(const Array2 &) const override { return Array1{0.0}; }
  void update_parameters(const Array1 &d) override { p[0] += d[0]; }
  const Array1 &parameters() const override { return p; }
  size_t num_parameters() const override { return 1; }
};

struct CountingSampler : MonteCarloSampler<HostWave> {
  HostWave wf;
  DataMap d;
  size_t moves = 0, k = 0;
  Status sample() override {
    d["Energy"].push_back(OperatorValue::scalar((double)k++));
    return Ok();
  }
  const DataMap &data() const override { return d; }
  Status move_state() override {
    ++moves;
    return Ok();
  }
  size_t num_observables() const override { return d.size(); }
  double acceptance() const override { return (double)moves; }
  std::vector<std::string> observable_names() const override { return {"Energy"}; }
  MonteCarloResult<HostWave> consume_result() override { return {wf, acceptance(), d}; }
  const HostWave &wave_function() const override { return wf; }
  HostWave &wave_function_mut() override { return wf; }
  void reseed_rng(const Seed &) override {}
  Seed generate_seed() override { return Seed{}; }
};

struct MeanStep : Optimizer {
  Result<Array1> compute_parameter_update(const Array1 &, const AverageMap &a, const DataMap &) override {
    return Array1{-0.1 * a.at("Energy").get_scalar().value()};
  }
};

static void test_runner_schedule() {
  EmptyLogger log;
  CHECK(panics([&] { Runner<CountingSampler>(CountingSampler{}, log).run(10, 6); }));
  Runner<CountingSampler> r(CountingSampler{}, log);
  auto res = r.run(30, 10);
  CHECK(res.is_ok());
  // 30 moves, the first block of 10 unsampled
  CHECK(res.value().acceptance == 30.0);
  CHECK(res.value().data.at("Energy").size() == 20);
  // VmcRunner on the host path: mean of 0..19 = 9.5 with the running mean,
  // blocking error over blocks of 5 of the 20 samples
  CountingSampler s;
  MeanStep opt;
  VmcRunner<CountingSampler, MeanStep> vmc(s, opt, log);
  auto out = vmc.run_optimization(1, 30, 10, 1, false);
  CHECK(out.is_ok());
  CHECK(std::fabs(out.value().energies[0] - 9.5) < 1e-12);
  CHECK(std::fabs(out.value().parameters[0] - (1.0 - 0.95)) < 1e-12);
  // blocks of 10 over 0..19: means 4.5 and 14.5; sqrt(<B^2> - <B>^2) / 1
  CHECK(std::fabs(out.value().energy_errors[0] - 5.0) < 1e-12);
}

// The stateful optimizers of optimizers.rs:32-179 over fabricated "SR sums"
// (host arithmetic only): n = 4, sum_e = 8, sum_o = (2, 4), sum_oe = (6, 4)
// give Ebar = 2 and g = 2 (sum_oe - Ebar sum_o) / n = (1, -2).
static void test_stateful_optimizers() {
  mole_sr_sums r{};
  r.n_params = 2;
  r.n = 4;
  r.sum_e = 8.0;
  r.sum_o[0] = 2.0;
  r.sum_o[1] = 4.0;
  r.sum_oe[0] = 6.0;
  r.sum_oe[1] = 4.0;
  DataMap raw;
  raw["SR sums"] = {OperatorValue::matrix(GpuSampler<GpuWaveFunction>::pack_sums(r))};
  const AverageMap avg;
  const Array1 pars{0.0, 0.0};
  // MomentumDescent: m -= 0.5 g; dp = 0.9 m
  MomentumDescent md(0.5, 0.9, 2);
  Array1 dp = md.compute_parameter_update(pars, avg, raw).value();
  CHECK(dp[0] == 0.9 * -0.5 && dp[1] == 0.9 * 1.0);
  MomentumDescent copy = md;  // Clone: an independent optimizer
  dp = md.compute_parameter_update(pars, avg, raw).value();
  CHECK(dp[0] == 0.9 * -1.0 && dp[1] == 0.9 * 2.0);
  CHECK(copy.compute_parameter_update(pars, avg, raw).value()[0] == dp[0]);
  CHECK(md.last_gradient()[0] == 1.0 && md.last_gradient()[1] == -2.0);
  // NesterovMomentum: m = 0.9 * 0 + 0.5 g; dp = -(0.9 * 0 + 1.9 m)
  NesterovMomentum nm(0.5, 0.9, 2);
  dp = nm.compute_parameter_update(pars, avg, raw).value();
  CHECK(dp[0] == -(1.9 * 0.5) && dp[1] == -(1.9 * -1.0));
  // the history must hold at least one pair; the parameter count must match
  CHECK(panics([] { OnlineLbfgs(0.1, 0, 2); }));
  OnlineLbfgs lb(0.1, 4, 3);
  CHECK(lb.compute_parameter_update(pars, avg, raw).is_err());
  CHECK(lb.state().history == 4 && lb.state().pars_prev[0] == 1e-5);
}

static int run_gpu(int iters, int walkers, int sweeps, int block, double box, double step, double eps,
                   uint64_t seed) {
  GpuSampler<GpuWaveFunction> sampler(GpuWaveFunction::hydrogen_molecule(1.0, 0.5, 0.3), (uint32_t)walkers, box,
                                      seed, (uint32_t)block);
  if (!sampler.ok()) {
    std::fprintf(stderr, "context: %s\n", sampler.error() ? sampler.error()->message.c_str() : "?");
    return 2;
  }
  StochasticReconfiguration sr(step, MOLE_SR_TEXTBOOK, eps);
  EmptyLogger log;
  VmcRunner<GpuSampler<GpuWaveFunction>, StochasticReconfiguration> vmc(sampler, sr, log);
  auto out = vmc.run_optimization((size_t)iters, (size_t)sweeps, (size_t)block, 1, false);
  if (out.is_err()) {
    std::fprintf(stderr, "%s: %s\n", error_name(out.error().kind), out.error().message.c_str());
    return 3;
  }
  const OptimizationResult &r = out.value();
  // a streaming Runner::run: Log::log once per sampled block with the block
  // means so far (mole_read_blocks)
  struct CountLog : Log {
    size_t calls = 0, last_blocks = 0;
    std::string log(const DataMap &d) override {
      ++calls;
      const auto it = d.find("Energy block means");
      last_blocks = it == d.end() ? 0 : it->second.size();
      return std::string();
    }
  } counter;
  GpuSampler<GpuWaveFunction> streaming(GpuWaveFunction::hydrogen_molecule(1.0, 0.5, 0.3), 256u, box, seed,
                                        (uint32_t)block, (uint32_t)(5 * block));
  Runner<GpuSampler<GpuWaveFunction>> runner(std::move(streaming), counter);
  auto mc = runner.run((size_t)(6 * block), (size_t)block);
  if (mc.is_err()) {
    std::fprintf(stderr, "streaming run: %s\n", mc.error().message.c_str());
    return 4;
  }
  const size_t n_bm = mc.value().data.at("Energy block means").size();
  // examples/custom_operator.rs on the device: the Gaussian trial under
  // HarmonicHamiltonian, Runner::run(1000, 1), box side 1
  GpuSampler<GpuSlaterJastrow> sho(GpuSlaterJastrow::gaussian_oscillator(1.0), 64u, 1.0, 0, 1u);
  const Status hs = sho.set_hamiltonian(operators::harmonic_hamiltonian(1.0).op);
  if (hs.is_err()) {
    std::fprintf(stderr, "set_hamiltonian: %s\n", hs.error().message.c_str());
    return 5;
  }
  Runner<GpuSampler<GpuSlaterJastrow>> sho_runner(std::move(sho), log);
  auto sho_res = sho_runner.run(1000, 1);
  if (sho_res.is_err()) return 6;
  const double sho_energy = sho_res.value().data.at("Energy").front().get_scalar().value();
  // DmcRunner on the hydrogen 1s function (exact: every energy -1/2)
  GpuDmcRunner dmc(GpuSlaterJastrow::hydrogenic(1), 128, -0.5, std::nullopt, 0.05, 2, Branching::SR);
  const auto de = dmc.diffuse(0.05, 40, 10, 1);
  const double dmc_energy = de.first.empty() ? 0.0 : de.first.back();
  bool dmc_panics = false;
  try {
    GpuDmcRunner bad(GpuSlaterJastrow::helium(1.6875, 0.5, 0.3), 16, -1e3, std::nullopt, 0.02, 2, Branching::SR);
    (void)bad.diffuse(1e4, 2, 1, 0);  // every weight underflows: WeightedChoice's panic
  } catch (const Panic &) {
    dmc_panics = true;
  }
  // single-configuration evaluation through the traits
  GpuWaveFunction he = GpuWaveFunction::helium(1.6875, 0.5, 0.3);
  Array2 x(2, 3);
  x(0, 0) = 0.3; x(0, 1) = -0.2; x(0, 2) = 0.5; x(1, 0) = -0.4; x(1, 1) = 0.1; x(1, 2) = -0.6;
  const double v = he.value(x).value(), lap = he.laplacian(x).value(), el = he.local_energy(x).value();
  const Array2 g = he.gradient(x).value();
  // the same through GpuSlaterJastrow (mole_eval of a Slater-Jastrow system)
  const GpuSlaterJastrow hs_sj = GpuSlaterJastrow::helium(1.6875, 0.5, 0.3);
  const auto sj_v = hs_sj.value(x), sj_l = hs_sj.laplacian(x), sj_e = hs_sj.local_energy(x);
  const auto sj_g = hs_sj.gradient(x);
  const auto sj_p = hs_sj.parameter_gradient(x);
  if (sj_v.is_err() || sj_l.is_err() || sj_e.is_err() || sj_g.is_err() || sj_p.is_err()) {
    std::fprintf(stderr, "GpuSlaterJastrow evaluation failed\n");
    return 7;
  }
  const GpuSlaterJastrow copy = hs_sj;  // a copy evaluates on its own context
  const bool copy_same = copy.value(x).value() == sj_v.value();
  std::printf("{\"params\": [%.17g, %.17g, %.17g], \"energies\": [", r.parameters[0], r.parameters[1], r.parameters[2]);
  for (size_t k = 0; k < r.energies.size(); ++k) std::printf("%s%.17g", k ? ", " : "", r.energies[k]);
  std::printf("], \"errors\": [");
  for (size_t k = 0; k < r.energy_errors.size(); ++k) std::printf("%s%.17g", k ? ", " : "", r.energy_errors[k]);
  std::printf("], \"he_value\": %.17g, \"he_laplacian\": %.17g, \"he_local_energy\": %.17g, \"he_grad\": [", v, lap, el);
  for (size_t k = 0; k < 6; ++k) std::printf("%s%.17g", k ? ", " : "", g.data[k]);
  std::printf("], \"sj_value\": %.17g, \"sj_laplacian\": %.17g, \"sj_local_energy\": %.17g, \"sj_copy_same\": %d, "
              "\"sj_grad\": [", sj_v.value(), sj_l.value(), sj_e.value(), copy_same ? 1 : 0);
  for (size_t k = 0; k < 6; ++k) std::printf("%s%.17g", k ? ", " : "", sj_g.value().data[k]);
  std::printf("], \"sj_pgrad\": [");
  for (size_t k = 0; k < sj_p.value().size(); ++k) std::printf("%s%.17g", k ? ", " : "", sj_p.value()[k]);
  std::printf("], \"stream_calls\": %zu, \"stream_blocks\": %zu, \"result_blocks\": %zu, \"sho_energy\": %.17g, "
              "\"dmc_energy\": %.17g, \"dmc_blocks\": %zu, \"dmc_panics\": %d}\n",
              counter.calls, counter.last_blocks, n_bm, sho_energy, dmc_energy, de.first.size(), dmc_panics ? 1 : 0);
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "--gpu") == 0) {
    if (argc != 10) {
      std::fprintf(stderr, "usage: host_tests --gpu iters walkers sweeps block box step eps seed\n");
      return 2;
    }
    return run_gpu(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                   std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]), std::strtoull(argv[9], nullptr, 10));
  }
  test_operator_value();
  test_status_mapping();
  test_runner_schedule();
  test_stateful_optimizers();
  if (g_fail) {
    std::fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  std::printf("host tests passed\n");
  return 0;
}
