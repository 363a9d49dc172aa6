// Runner (src/montecarlo/src/montecarlo.rs:15-50) and VmcRunner
// (src/vmc/src/vmc.rs:21-180) over the host traits.
//
// Runner::run: blocks of block_size sweeps, the first block discarded; a
// FusedSampler (the device sampler) runs the same schedule in launches of
// many sweeps.  VmcRunner::run_optimization: per iteration, the sampler's
// observables, their averages and blocking errors, the optimizer's step
// and update_parameters.  With a FusedSampler the walkers persist across
// iterations and the averages come from the device reductions; otherwise
// nworkers copies of the sampler run from the same state with fresh seeds
// (vmc.rs:56-76) and the per-sample data are concatenated and reduced as
// vmc.rs:108-179 does (running means, blocking errors).
#pragma once

#include <cmath>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "gpu.hpp"
#include "traits.hpp"

namespace mole {

template <typename S>
class Runner {
 public:
  Runner(S sampler, Log &logger) : sampler_(std::move(sampler)), logger_(logger) {}

  Result<MonteCarloResult<typename S::WaveFunc>> run(size_t steps, size_t block_size) {
    if (!(steps >= 2 * block_size)) throw Panic("assertion failed: steps >= 2 * block_size");
    if (auto *f = dynamic_cast<FusedSampler *>(&sampler_)) {
      if (f->streams_blocks()) {  // Log::log after every sampled block, as the reference prints
        const Status st = f->run_fused_streaming(steps, block_size, [this] {
          const std::string out = logger_.log(sampler_.data());
          if (!out.empty()) std::printf("%s\n", out.c_str());
        });
        if (st.is_err()) return st.error();
        return sampler_.consume_result();
      }
      const Status st = f->run_fused(steps, block_size);
      if (st.is_err()) return st.error();
      const std::string out = logger_.log(sampler_.data());
      if (!out.empty()) std::printf("%s\n", out.c_str());
      return sampler_.consume_result();
    }
    const size_t blocks = steps / block_size;
    std::string output;
    for (size_t block_nr = 0; block_nr < blocks; ++block_nr) {
      for (size_t k = 0; k < block_size; ++k) {
        Status st = sampler_.move_state();
        if (st.is_err()) return st.error();
        if (block_nr > 0) {  // discard the first block for equilibration
          st = sampler_.sample();
          if (st.is_err()) return st.error();
          output = logger_.log(sampler_.data());
        }
      }
      if (!output.empty()) std::printf("%s\n", output.c_str());
    }
    return sampler_.consume_result();
  }

  const DataMap &data() const { return sampler_.data(); }
  S &sampler() { return sampler_; }

 private:
  S sampler_;
  Log &logger_;
};

struct OptimizationResult {
  Array1 parameters;
  Array1 energies;
  Array1 energy_errors;
  Array1 acceptances;
};

template <typename S, typename O>
class VmcRunner {
 public:
  VmcRunner(S &sampler, O &optimizer, Log &logger) : sampler_(sampler), optimizer_(optimizer), logger_(logger) {}

  // iters SR/SD iterations of total_samples sweeps per worker set
  Result<OptimizationResult> run_optimization(size_t iters, size_t total_samples, size_t block_size,
                                              size_t nworkers, bool print = true) {
    OptimizationResult res;
    const size_t steps = total_samples / nworkers;
    for (size_t it = 0; it < iters; ++it) {
      AverageMap averages;
      DataMap data;
      double acceptance = 0.0;
      if (auto *f = dynamic_cast<FusedSampler *>(&sampler_)) {
        if (nworkers != 1) throw Panic("a device sampler holds its walkers in one context: nworkers must be 1");
        const Status st = f->run_fused(steps, block_size);
        if (st.is_err()) return st.error();
        data = sampler_.data();
        acceptance = sampler_.acceptance();
        const auto e = data.find("Energy"), er = data.find("Energy error");
        if (e == data.end() || er == data.end()) return Error{ErrorKind::DataAccessError, "no energy samples"};
        averages["Energy"] = e->second.back();
        res.energy_errors.push_back(er->second.back().get_scalar().value());
      } else {
        std::vector<MonteCarloResult<typename S::WaveFunc>> results;
        for (size_t w = 0; w < nworkers; ++w) {
          S worker = sampler_;
          worker.reseed_rng(sampler_.generate_seed());
          EmptyLogger quiet;
          Runner<S> r(worker, w == 0 ? logger_ : quiet);
          auto out = r.run(steps, block_size);
          if (out.is_err()) return out.error();
          results.push_back(out.value());
        }
        for (auto &r : results) {  // concatenate_worker_data (vmc.rs:108-130)
          acceptance += r.acceptance;
          for (auto &kv : r.data) data[kv.first].insert(data[kv.first].end(), kv.second.begin(), kv.second.end());
        }
        AverageMap errors;
        process_monte_carlo_results(data, block_size, averages, errors);
        res.energy_errors.push_back(errors.at("Energy").get_scalar().value());
      }
      res.energies.push_back(averages.at("Energy").get_scalar().value());
      res.acceptances.push_back(acceptance);
      auto dp = optimizer_.compute_parameter_update(sampler_.wave_function().parameters(), averages, data);
      if (dp.is_err()) return dp.error();
      sampler_.wave_function_mut().update_parameters(dp.value());
      if (print)
        std::printf("Energy:      %.8f +/- %.9f    accept: %.8f\n", res.energies.back(), res.energy_errors.back(),
                    acceptance / (double)total_samples);
    }
    res.parameters = sampler_.wave_function().parameters();
    return res;
  }

  // vmc.rs:174-179: running mean a += (b - a)/(n + 1)
  static OperatorValue mean(const std::vector<OperatorValue> &v, size_t first = 0, size_t last = size_t(-1)) {
    if (last > v.size()) last = v.size();
    OperatorValue a = OperatorValue::scalar(0.0);
    for (size_t n = first; n < last; ++n) a = a + (v[n] - a) / OperatorValue::scalar((double)(n - first + 1));
    return a;
  }

  // vmc.rs:135-172: averages and blocking errors of every observable
  static void process_monte_carlo_results(const DataMap &mc, size_t block_size, AverageMap &averages,
                                          AverageMap &errors) {
    for (const auto &kv : mc) averages.emplace(kv.first, mean(kv.second));
    for (const auto &kv : mc) {
      const OperatorValue &m = averages.at(kv.first);
      std::vector<OperatorValue> block_means;
      for (size_t b = 0; b < kv.second.size(); b += block_size)
        block_means.push_back(mean(kv.second, b, b + block_size));
      std::vector<OperatorValue> sq;
      for (const auto &x : block_means) sq.push_back(x * x);
      const OperatorValue bms = mean(sq);
      const double nb = (double)block_means.size();
      errors.emplace(kv.first, ((bms - m * m) / OperatorValue::scalar(nb - 1.0)).map([](double x) { return std::sqrt(x); }));
    }
  }

 private:
  S &sampler_;
  O &optimizer_;
  Log &logger_;
};

}  // namespace mole
