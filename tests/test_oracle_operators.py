"""Pins of the oracle's Gaussian (GTO) basis functions and local operators
(oracle/sj.c: basis kinds, orc_sj_operator):

* GTO and mixed STO/GTO wavefunctions: the drift (grad ln psi) and the
  KINETIC term (-1/2 lap psi / psi) against central finite differences of
  the oracle's own psi (LU determinants, no derivative code involved).
* Each potential term against a numpy evaluation on the configuration.
* The ElectronicHamiltonian built from terms equals the built-in local
  energy; the oscillator of examples/custom_operator.rs (Gaussian trial
  exp(-(r/a)^2), a = sqrt(2/omega), T + 1/2 omega^2 r^2) has E_L = 3 omega/2
  at every configuration (the reference asserts the mean to 1e-15 and the
  spread below 1e-15, custom_operator.rs:135-136); a VMC run of it samples
  that value at every step.
"""
import numpy as np
import pytest

from mole_amd import systems as S
from oracle import sj_ffi as SJ

T = SJ


def gto_atom():
    # He-like: two electrons in a GTO/STO mixture with the Pade Jastrow on
    return S.SJSystem([(0.0, 0.0, 0.0)], [2], 1, 1, [0.8, 1.7, 2.3], [0, 0, 0], [[0.6, 0.3, 0.5]],
                      (0.5, 0.25, 1.0, 0.3, 1.0), name="He GTO", kind=[1, 0, 1])


def gto_molecule():
    # 3 electrons on two centres, all GTO, two orbitals
    return S.SJSystem([(0.0, 0.0, -0.7), (0.0, 0.0, 0.7)], [2, 1], 2, 1, [0.5, 1.2, 0.7, 1.5], [0, 0, 1, 1],
                      [[0.7, 0.2, 0.4, 0.1], [0.1, -0.6, 0.3, 0.8]], (0.5, 0.25, 1.0, 0.3, 1.0),
                      name="GTO molecule", kind=[1, 1, 1, 1])


def fd_lap_and_grad(wv, x, h=1e-4):
    f0 = SJ.value(wv, x)
    g = np.zeros_like(x)
    lap = 0.0
    for k in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fp, fm = SJ.value(wv, xp), SJ.value(wv, xm)
        g[k] = (fp - fm) / (2 * h) / f0
        lap += (fp - 2 * f0 + fm) / (h * h)
    return lap / f0, g


@pytest.mark.parametrize("make", [gto_atom, gto_molecule])
def test_gto_derivatives_against_finite_differences(make):
    sj = make()
    wv = SJ.SjWave(sj)
    rng = np.random.default_rng(3)
    for _ in range(5):
        x = rng.normal(0, 0.9, 3 * sj.n_electrons)
        lap, g = fd_lap_and_grad(wv, x)
        kin = SJ.operator_value(wv, x, [(T.TERM_KINETIC, 1.0)])
        assert abs(kin - (-0.5 * lap)) < 2e-5 * max(1.0, abs(kin))
        for i in range(sj.n_electrons):
            assert np.allclose(SJ.drift(wv, x, i), g[3 * i:3 * i + 3], rtol=1e-6, atol=1e-6)


def test_potential_terms_against_numpy():
    sj = gto_molecule()
    wv = SJ.SjWave(sj)
    rng = np.random.default_rng(4)
    nuc, Z = sj.nuclei, sj.charges.astype(float)
    for _ in range(5):
        x = rng.normal(0, 1.0, 3 * sj.n_electrons)
        r = x.reshape(-1, 3)
        vee = sum(1.0 / np.linalg.norm(r[i] - r[j]) for i in range(len(r)) for j in range(i + 1, len(r)))
        ven = -sum(Z[I] / np.linalg.norm(r[i] - nuc[I]) for I in range(len(nuc)) for i in range(len(r)))
        vnn = Z[0] * Z[1] / np.linalg.norm(nuc[0] - nuc[1])
        c = np.array([0.1, -0.2, 0.3])
        u = np.array([0.0, 0.6, 0.8])
        assert abs(SJ.operator_value(wv, x, [(T.TERM_ELECTRONIC_POTENTIAL, 1.0)]) - vee) < 1e-13 * abs(vee)
        assert abs(SJ.operator_value(wv, x, [(T.TERM_IONIC_POTENTIAL, 1.0)]) - (ven + vnn)) < 1e-13 * abs(ven)
        h = SJ.operator_value(wv, x, [(T.TERM_HARMONIC, 2.0, 1.3, *c)])
        assert abs(h - 2.0 * 0.5 * 1.3**2 * np.sum((r - c) ** 2)) < 1e-13 * abs(h)
        for k in (-2, -1, 1, 3):
            v = SJ.operator_value(wv, x, [(T.TERM_RADIAL_POWER, 1.0, k, *c)])
            assert abs(v - np.sum(np.linalg.norm(r - c, axis=1) ** k)) < 1e-12 * abs(v)
        d = SJ.operator_value(wv, x, [(T.TERM_DIPOLE, -1.5, *c, *u)])
        assert abs(d - (-1.5) * np.sum((r - c) @ u)) < 1e-13 * max(1.0, abs(d))


@pytest.mark.parametrize("make", [gto_atom, gto_molecule, lambda: S.atom(3)])
def test_electronic_hamiltonian_as_terms_equals_local_energy(make):
    sj = make()
    wv = SJ.SjWave(sj)
    rng = np.random.default_rng(5)
    terms = [(T.TERM_KINETIC, 1.0), (T.TERM_IONIC_POTENTIAL, 1.0), (T.TERM_ELECTRONIC_POTENTIAL, 1.0)]
    for _ in range(5):
        x = rng.normal(0, 1.0, 3 * sj.n_electrons)
        e, _ = SJ.local_energy(wv, x)
        assert abs(SJ.operator_value(wv, x, terms) - e) < 1e-12 * max(1.0, abs(e))


@pytest.mark.parametrize("omega", [1.0, 0.5, 2.0])
def test_oscillator_zero_variance(omega):
    sj = S.harmonic_oscillator(omega)
    wv = SJ.SjWave(sj)
    h = [(T.TERM_KINETIC, 1.0), (T.TERM_HARMONIC, 1.0, omega, 0.0, 0.0, 0.0)]
    rng = np.random.default_rng(6)
    for _ in range(20):
        x = rng.normal(0, 1.5, 3)
        assert abs(SJ.operator_value(wv, x, h) - 1.5 * omega) < 1e-13 * max(1.0, float(x @ x))
    # Runner::run of the example (box side 1, 1000 steps) samples 3 omega/2 every step
    out = SJ.run(wv, seed=0, walkers=8, steps=1000, block_size=1, box_side=1.0, hamiltonian=h)
    el = out["e_local"]
    assert np.all(np.abs(el - 1.5 * omega) < 1e-12)
    # a wrong width (a != sqrt(2/omega)) is not zero-variance: the pin discriminates
    wv2 = SJ.SjWave(S.harmonic_oscillator(omega, a=1.1 * np.sqrt(2.0 / omega)))
    out2 = SJ.run(wv2, seed=0, walkers=8, steps=200, block_size=1, box_side=1.0, hamiltonian=h)
    assert np.std(out2["e_local"]) > 1e-3
