// Slater-Jastrow systems and device local operators for the host layer.
//
// GpuSlaterJastrow is the wave function of a Slater-Jastrow system as plain
// data (nuclei, spins, s-type STO/GTO basis, orbital coefficients, the five
// Jastrow parameters b_anti, b_para, c, a, d, and optionally the exponent
// scale): the WaveFunction and Optimize traits (src/wavefunction_traits/src/
// lib.rs:26-28, src/optimize/src/traits.rs:8-16) for GpuSampler and
// GpuDmcRunner, and the per-configuration Function/Differentiate
// evaluations through mole_eval (FP64, on a one-walker context of its own
// that copies do not share).
//
// The operator builders make the terms of mole_operator: KineticEnergy,
// IonicPotential, ElectronicPotential (src/operator/src/operator.rs:12-121),
// the harmonic potential of examples/custom_operator.rs:48-61, radial powers
// and a dipole; `+` concatenates terms into one operator (a sum, as the
// reference's Hamiltonians add their parts' OperatorValues).
#pragma once

#include <cstring>
#include <memory>
#include <vector>

#include "errors.hpp"
#include "gpu.hpp"
#include "traits.hpp"

namespace mole {

class GpuSlaterJastrow : public Function, public Differentiate, public WaveFunction, public Optimize {
 public:
  GpuSlaterJastrow(std::vector<double> nuclei_xyz, std::vector<int32_t> charges, uint32_t n_up, uint32_t n_down,
                   std::vector<double> zeta, std::vector<uint32_t> center, std::vector<double> coef,
                   Array1 jastrow, std::vector<uint32_t> kind = {}, uint32_t hamiltonian = MOLE_H_ELECTRONIC)
      : nuc_(std::move(nuclei_xyz)), z_(std::move(charges)), n_up_(n_up), n_down_(n_down),
        zeta_(std::move(zeta)), center_(std::move(center)), coef_(std::move(coef)), params_(std::move(jastrow)),
        kind_(std::move(kind)), hamiltonian_(hamiltonian) {
    if (kind_.empty()) kind_.assign(zeta_.size(), MOLE_BASIS_STO);
  }

  // one electron in exp(-(r/a)^2), a = sqrt(2/omega), about a charge-0 centre:
  // the trial function of examples/custom_operator.rs:64-100
  static GpuSlaterJastrow gaussian_oscillator(double omega) {
    const double a2 = 2.0 / omega;
    return GpuSlaterJastrow({0.0, 0.0, 0.0}, {0}, 1, 0, {1.0 / a2}, {0}, {1.0}, {0, 0, 0, 0, 0},
                            {MOLE_BASIS_GTO});
  }
  // hydrogen-like 1s exp(-Z r): an exact eigenfunction, E_L = -Z^2/2
  static GpuSlaterJastrow hydrogenic(int32_t Z) {
    return GpuSlaterJastrow({0.0, 0.0, 0.0}, {Z}, 1, 0, {(double)Z}, {0}, {1.0}, {0, 0, 0, 0, 0});
  }
  // He with exp(-zeta r) orbitals and the e-e Pade factor b r/(1 + c r)
  static GpuSlaterJastrow helium(double zeta, double b, double c) {
    return GpuSlaterJastrow({0.0, 0.0, 0.0}, {2}, 1, 1, {zeta}, {0}, {1.0}, {b, 0.0, c, 0.0, 1.0});
  }

  mole_system_desc desc(uint32_t walkers) const {
    mole_system_desc d{};
    d.wf_kind = MOLE_WF_SLATER_JASTROW;
    d.hamiltonian = hamiltonian_;
    d.n_up = n_up_;
    d.n_down = n_down_;
    d.n_nuclei = (uint32_t)z_.size();
    d.nuclei_xyz = nuc_.data();
    d.nuclei_charge = z_.data();
    d.n_params = (uint32_t)params_.size();
    d.params = params_.data();
    d.n_walkers = walkers;
    d.n_basis = (uint32_t)zeta_.size();
    d.basis_exponent = zeta_.data();
    d.basis_center = center_.data();
    d.n_orbitals = (uint32_t)(coef_.size() / zeta_.size());
    d.orbital_coeff = coef_.data();
    d.basis_kind = kind_.data();
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
    Array2 g(num_electrons(), 3);
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
  // the system's Hamiltonian acting on psi, divided by psi (operator.rs:177-183)
  Result<double> local_energy(const Array2 &cfg) const {
    double e = 0.0;
    const Status st = eval(cfg, nullptr, nullptr, nullptr, &e, nullptr);
    if (st.is_err()) return st.error();
    return e;
  }
  size_t num_electrons() const override { return n_up_ + n_down_; }
  // Optimize::parameter_gradient: d psi / dp
  Result<Array1> parameter_gradient(const Array2 &cfg) const override {
    Array1 g(params_.size());
    const Status st = eval(cfg, nullptr, nullptr, nullptr, nullptr, g.data());
    if (st.is_err()) return st.error();
    return g;
  }
  void update_parameters(const Array1 &deltap) override {
    for (size_t k = 0; k < params_.size() && k < deltap.size(); ++k) params_[k] += deltap[k];
    cache_.ctx.reset();
  }
  const Array1 &parameters() const override { return params_; }
  size_t num_parameters() const override { return params_.size(); }

 private:
  // the evaluation context: created on first use, dropped when the
  // parameters change, never shared by copies
  struct EvalCache {
    mutable std::shared_ptr<gpu::Context> ctx;
    EvalCache() = default;
    EvalCache(const EvalCache &) {}
    EvalCache &operator=(const EvalCache &) {
      ctx.reset();
      return *this;
    }
  };
  Status eval(const Array2 &cfg, double *v, double *g, double *l, double *e, double *pg) const {
    if (cfg.rows != num_electrons() || cfg.cols != 3)
      return Error{ErrorKind::ShapeError, "expected an N x 3 configuration"};
    if (!cache_.ctx) {
      cache_.ctx = std::make_shared<gpu::Context>(std::vector<mole_system_desc>{desc(1)}, gpu::run_desc(0, 0.5, 1));
      if (!cache_.ctx->ok()) {
        const Error err = *cache_.ctx->error();
        cache_.ctx.reset();
        return err;
      }
    }
    return status(mole_eval(cache_.ctx->get(), 0, 1, cfg.data.data(), v, g, l, e, pg));
  }
  std::vector<double> nuc_;
  std::vector<int32_t> z_;
  uint32_t n_up_, n_down_;
  std::vector<double> zeta_;
  std::vector<uint32_t> center_;
  std::vector<double> coef_;
  Array1 params_;
  std::vector<uint32_t> kind_;
  uint32_t hamiltonian_;
  EvalCache cache_;
};

// ---- operators as device terms ----
struct DeviceOperator {
  mole_operator op{};
  DeviceOperator operator+(const DeviceOperator &o) const {
    if (op.n_terms + o.op.n_terms > MOLE_OPERATOR_MAX_TERMS) throw Panic("an operator holds at most 4 terms");
    DeviceOperator r = *this;
    for (uint32_t k = 0; k < o.op.n_terms; ++k) r.op.term[r.op.n_terms++] = o.op.term[k];
    return r;
  }
};

namespace operators {
inline DeviceOperator term(uint32_t kind, double coef, std::initializer_list<double> param = {}) {
  DeviceOperator d;
  d.op.n_terms = 1;
  d.op.term[0].kind = kind;
  d.op.term[0].coef = coef;
  size_t q = 0;
  for (double v : param) d.op.term[0].param[q++] = v;
  return d;
}
inline DeviceOperator kinetic_energy() { return term(MOLE_TERM_KINETIC, 1.0); }
inline DeviceOperator ionic_potential() { return term(MOLE_TERM_IONIC_POTENTIAL, 1.0); }
inline DeviceOperator electronic_potential() { return term(MOLE_TERM_ELECTRONIC_POTENTIAL, 1.0); }
inline DeviceOperator electronic_hamiltonian() { return kinetic_energy() + ionic_potential() + electronic_potential(); }
inline DeviceOperator ionic_hamiltonian() { return kinetic_energy() + ionic_potential(); }
// 1/2 omega^2 |x - c|^2 over all electrons
inline DeviceOperator harmonic_potential(double omega, double cx = 0.0, double cy = 0.0, double cz = 0.0) {
  return term(MOLE_TERM_HARMONIC, 1.0, {omega, cx, cy, cz});
}
// the HarmonicHamiltonian of examples/custom_operator.rs:30-62
inline DeviceOperator harmonic_hamiltonian(double omega) { return kinetic_energy() + harmonic_potential(omega); }
}  // namespace operators

}  // namespace mole
