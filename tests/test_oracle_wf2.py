"""Pins the oracle's two-electron wavefunctions, potentials and local energy:
finite differences for gradient/laplacian/parameter gradient, the He closed
form of E_L, zero-variance eigenfunctions and the reference's potentials."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle_ffi as O


def _cfg(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, O.dptr(a)


def value(wf, x):
    a, p = _cfg(x)
    return O.lib().orc_wf2_value(C.byref(wf), p)


def grad(wf, x):
    a, p = _cfg(x)
    g = np.zeros(6)
    O.lib().orc_wf2_gradient(C.byref(wf), p, O.dptr(g))
    return g


def lap(wf, x):
    a, p = _cfg(x)
    return O.lib().orc_wf2_laplacian(C.byref(wf), p)


def el(wf, x):
    a, p = _cfg(x)
    return O.lib().orc_wf2_local_energy(C.byref(wf), p)


def pgrad(wf, x):
    a, p = _cfg(x)
    g = np.zeros(3)
    O.lib().orc_wf2_parameter_gradient(C.byref(wf), p, O.dptr(g))
    return g


def random_cfgs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.5, 1.5, size=(n, 6))


WFS = [O.helium(), O.helium(2.0, 0.25, 1.1), O.h2(), O.h2(1.3, 0.4, 0.7)]


@pytest.mark.parametrize("wf", WFS)
def test_gradient_matches_central_differences(wf):
    h = 1e-6
    for x in random_cfgs(20, 1):
        g = grad(wf, x)
        for k in range(6):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd = (value(wf, xp) - value(wf, xm)) / (2 * h)
            assert abs(g[k] - fd) <= 1e-7 * max(1.0, abs(fd)) + 1e-8 * value(wf, x)


@pytest.mark.parametrize("wf", WFS)
def test_laplacian_matches_finite_differences(wf):
    h = 1e-4
    for x in random_cfgs(20, 2):
        f0 = value(wf, x)
        fd = 0.0
        for k in range(6):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd += (value(wf, xp) - 2 * f0 + value(wf, xm)) / h**2
        assert abs(lap(wf, x) - fd) <= 2e-6 * max(abs(fd), f0)


@pytest.mark.parametrize("wf", WFS)
def test_parameter_gradient_matches_finite_differences(wf):
    h = 1e-6
    for x in random_cfgs(20, 3):
        g = pgrad(wf, x)
        for k in range(3):
            w2 = O.Wf2.from_buffer_copy(wf)
            w2.p[k] = wf.p[k] + h
            fp = value(w2, x)
            w2.p[k] = wf.p[k] - h
            fm = value(w2, x)
            fd = (fp - fm) / (2 * h)
            assert abs(g[k] - fd) <= 1e-7 * max(1.0, abs(fd))


def he_closed_form(zeta, b, c, x):
    """Appendix-B style closed form of E_L for exp(-zeta(r1+r2) + u(r12))."""
    r1v, r2v = x[:3], x[3:]
    r1, r2 = np.linalg.norm(r1v), np.linalg.norm(r2v)
    d = r1v - r2v
    r12 = np.linalg.norm(d)
    den = 1 + c * r12
    up = b / den**2
    upp = -2 * b * c / den**3
    rh12 = d / r12
    return ((zeta - 2) * (1 / r1 + 1 / r2) + 1 / r12 - zeta**2 - upp - 2 * up / r12 - up**2
            + zeta * up * np.dot(r1v / r1 - r2v / r2, rh12))


@pytest.mark.parametrize("params", [(1.6875, 0.5, 0.3), (2.0, 0.25, 1.0), (1.2, 0.0, 0.5)])
def test_helium_local_energy_closed_form(params):
    wf = O.helium(*params)
    for x in random_cfgs(200, 4):
        want = he_closed_form(*params, x)
        got = el(wf, x)
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want))


def test_helium_zero_variance_ionic_hamiltonian():
    # zeta = Z = 2, no Jastrow, IonicHamiltonian (operator.rs:130-149): E_L == -4
    wf = O.helium(2.0, 0.0, 0.0, hamiltonian=O.H_IONIC)
    for x in random_cfgs(200, 5):
        assert abs(el(wf, x) + 4.0) <= 1e-12


def test_ionic_potential_and_nuclear_repulsion():
    L = O.lib()
    wf = O.h2()
    x = np.array([0.1, -0.2, 0.3, -0.4, 0.5, 0.6])
    A, B = np.array([0, 0, -0.7]), np.array([0, 0, 0.7])
    want = -sum(1 / np.linalg.norm(x[3 * e:3 * e + 3] - R) for R in (A, B) for e in range(2)) + 1 / 1.4
    assert abs(L.orc_ionic_potential(C.byref(wf), O.dptr(x), 2) - want) <= 1e-14
    assert abs(L.orc_electronic_potential(O.dptr(x), 2) - 1 / np.linalg.norm(x[:3] - x[3:])) <= 1e-15


def test_h2_local_energy_is_kinetic_plus_potential():
    # E_L = -1/2 lap(psi)/psi + V, with lap from finite differences (independent path)
    wf = O.h2()
    h = 1e-4
    for x in random_cfgs(20, 6):
        f0 = value(wf, x)
        fd = 0.0
        for k in range(6):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd += (value(wf, xp) - 2 * f0 + value(wf, xm)) / h**2
        A, B = np.array([0, 0, -0.7]), np.array([0, 0, 0.7])
        V = (-sum(1 / np.linalg.norm(x[3 * e:3 * e + 3] - R) for R in (A, B) for e in range(2))
             + 1 / 1.4 + 1 / np.linalg.norm(x[:3] - x[3:]))
        want = -0.5 * fd / f0 + V
        assert abs(el(wf, x) - want) <= 1e-5 * max(1.0, abs(want))


def test_value_spelling_helium():
    wf = O.helium(1.6875, 0.5, 0.3)
    x = np.array([0.3, -0.1, 0.2, -0.5, 0.4, 0.1])
    r1 = np.linalg.norm(x[:3])
    r2 = np.linalg.norm(x[3:])
    r12 = np.linalg.norm(x[:3] - x[3:])
    want = math.exp(-1.6875 * (r1 + r2) + 0.5 * r12 / (1 + 0.3 * r12))
    assert abs(value(wf, x) - want) <= 4e-16 * want
