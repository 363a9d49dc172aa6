// The reference's traits, as abstract C++ interfaces (same names and
// argument meaning):
//   Function, Differentiate, WaveFunction   src/wavefunction_traits/src/lib.rs:7-24
//   Optimize, Optimizer                     src/optimize/src/traits.rs:8-25
//   LocalOperator                           src/operator/src/traits.rs:263-265
//   MonteCarloSampler, Log, MonteCarloResult src/montecarlo/src/traits.rs:5-47
// Configurations are Array2 (rows = electrons, 3 columns), owned by the
// caller and borrowed for the call, as &Array2<f64> in the reference.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "errors.hpp"
#include "operator_value.hpp"

namespace mole {

struct Array2 {
  size_t rows = 0, cols = 3;
  std::vector<double> data;  // row-major
  Array2() = default;
  Array2(size_t r, size_t c) : rows(r), cols(c), data(r * c, 0.0) {}
  double &operator()(size_t i, size_t j) { return data[i * cols + j]; }
  double operator()(size_t i, size_t j) const { return data[i * cols + j]; }
};
using Array1 = std::vector<double>;

struct Function {
  virtual ~Function() = default;
  virtual Result<double> value(const Array2 &cfg) const = 0;
};

struct Differentiate {
  virtual ~Differentiate() = default;
  virtual Result<Array2> gradient(const Array2 &cfg) const = 0;
  virtual Result<double> laplacian(const Array2 &cfg) const = 0;
};

struct WaveFunction {
  virtual ~WaveFunction() = default;
  virtual size_t num_electrons() const = 0;
};

struct Optimize {
  virtual ~Optimize() = default;
  virtual Result<Array1> parameter_gradient(const Array2 &cfg) const = 0;
  virtual void update_parameters(const Array1 &deltap) = 0;
  virtual const Array1 &parameters() const = 0;
  virtual size_t num_parameters() const = 0;
};

struct Optimizer {
  virtual ~Optimizer() = default;
  virtual Result<Array1> compute_parameter_update(const Array1 &pars, const AverageMap &averages,
                                                  const DataMap &raw_data) = 0;
};

template <typename T>
struct LocalOperator {
  virtual ~LocalOperator() = default;
  virtual Result<OperatorValue> act_on(const T &wf, const Array2 &cfg) const = 0;
};

template <typename T>
struct MonteCarloResult {
  T wave_function;
  double acceptance;
  DataMap data;
};

using Seed = std::array<uint8_t, 32>;

template <typename WF>
struct MonteCarloSampler {
  using WaveFunc = WF;
  virtual ~MonteCarloSampler() = default;
  virtual Status sample() = 0;
  virtual const DataMap &data() const = 0;
  virtual Status move_state() = 0;
  virtual size_t num_observables() const = 0;
  virtual double acceptance() const = 0;
  virtual std::vector<std::string> observable_names() const = 0;
  virtual MonteCarloResult<WF> consume_result() = 0;
  virtual const WF &wave_function() const = 0;
  virtual WF &wave_function_mut() = 0;
  virtual void reseed_rng(const Seed &seed) = 0;
  virtual Seed generate_seed() = 0;
};

struct Log {
  virtual ~Log() = default;
  virtual std::string log(const DataMap &data) = 0;
};

struct EmptyLogger : Log {
  std::string log(const DataMap &) override { return std::string(); }
};

}  // namespace mole
