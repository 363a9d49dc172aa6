// DmcRunner (src/dmc/src/dmc.rs:16-203) over the device DMC of the C ABI:
// the guiding wave function is a GpuSlaterJastrow, the MetropolisDiffuse a
// MOLE_PROPOSAL_DIFFUSE context of time step tau, the branching SRBrancher
// or SimpleBranching (src/dmc/src/branching.rs:7-91), the Hamiltonian the
// built-in one of the system or a device operator.  diffuse returns the
// reference's (energies, errors) and panics where the reference unwraps.
#pragma once

#include <optional>
#include <utility>
#include <vector>

#include "gpu.hpp"
#include "slater_jastrow.hpp"

namespace mole {

enum class Branching : uint32_t { SR = MOLE_BRANCH_SR, Simple = MOLE_BRANCH_SIMPLE };

class GpuDmcRunner {
 public:
  // DmcRunner::new(guiding_wave_function, num_walkers, reference_energy,
  // hamiltonian, metropolis, branching) (dmc.rs:45-73)
  GpuDmcRunner(GpuSlaterJastrow guide, size_t num_walkers, double reference_energy,
               std::optional<DeviceOperator> hamiltonian, double tau, uint64_t seed, Branching branching,
               uint32_t precision = MOLE_PREC_FP64, uint32_t lu_period = 0)
      : guide_(std::move(guide)), n_walkers_(num_walkers) {
    const mole_system_desc d = guide_.desc((uint32_t)num_walkers);
    mole_run_desc r = gpu::run_desc(seed, tau, 1);
    r.proposal = MOLE_PROPOSAL_DIFFUSE;
    r.precision = precision;
    r.lu_period = lu_period;
    ctx_ = std::make_shared<gpu::Context>(std::vector<mole_system_desc>{d}, r);
    if (!ctx_->ok()) throw Panic("DmcRunner::new: " + ctx_->error()->message);
    if (hamiltonian) check(mole_set_hamiltonian(ctx_->get(), 0, &hamiltonian->op));
    check(mole_dmc_init(ctx_->get(), &reference_energy, (uint32_t)branching));
  }

  // DmcRunner::diffuse (dmc.rs:75-153)
  std::pair<std::vector<double>, std::vector<double>> diffuse(double time_step, size_t num_iterations,
                                                              size_t block_size, size_t num_eq_blocks) {
    check(mole_dmc_diffuse(ctx_->get(), time_step, (uint32_t)num_iterations, (uint32_t)block_size,
                           (uint32_t)num_eq_blocks));
    size_t n = 0;
    double eref = 0.0;
    check(mole_dmc_read(ctx_->get(), 0, nullptr, nullptr, 0, &n, &eref));
    std::vector<double> e(n), err(n);
    if (n) check(mole_dmc_read(ctx_->get(), 0, e.data(), err.data(), n, &n, &eref));
    return {e, err};
  }

  double reference_energy() const {
    size_t n = 0;
    double eref = 0.0;
    check(mole_dmc_read(ctx_->get(), 0, nullptr, nullptr, 0, &n, &eref));
    return eref;
  }
  // the walkers' weights (DmcRunner::walkers, dmc.rs:27)
  std::vector<double> weights() const {
    std::vector<double> w(walkers());
    check(mole_dmc_read_walkers(ctx_->get(), 0, w.data(), nullptr, w.size()));
    return w;
  }
  size_t walkers() const { return n_walkers_; }
  const GpuSlaterJastrow &guiding_wave_function() const { return guide_; }

 private:
  static void check(mole_status s) {
    const Status st = status(s);
    if (st.is_err()) throw Panic(std::string("called unwrap on an Err: ") + st.error().message);
  }
  GpuSlaterJastrow guide_;
  size_t n_walkers_;
  std::shared_ptr<gpu::Context> ctx_;
};

}  // namespace mole
