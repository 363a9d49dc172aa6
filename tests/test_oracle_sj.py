"""Pins the oracle's Slater-Jastrow restatement: LU determinants against the
Leibniz expansion, Gauss-Jordan inverses against A B = I, the local energy
against finite differences and zero-variance eigenfunctions (1s and 2p), the
Jastrow parameter derivatives against finite differences, the Sherman-Morrison
runner against the from-scratch (reference-faithful) runner, the pivot
determinants and the carried log|det| against LU and Leibniz, and the
singular-pivot guard on planted singular configurations."""
import ctypes as C

import numpy as np
import pytest

from mole_amd import systems as S
from oracle import oracle_ffi as OF
from oracle import sj_ffi as SJ


def cfgs(N, n, seed, scale=1.2):
    return np.random.default_rng(seed).normal(scale=scale, size=(n, 3 * N))


def basis_numpy(system, xe):
    """basis values at the point xe, written from the definitions (numpy)"""
    d = xe[None, :] - system.nuclei[system.center]
    r = np.linalg.norm(d, axis=1)
    kind = np.zeros(system.n_basis, int) if system.kind is None else system.kind
    v = np.where(kind == S.GTO, np.exp(-system.zeta * r * r), np.exp(-system.zeta * r))
    for q in range(3):
        v = np.where(kind == S.STO_PX + q, d[:, q] * np.exp(-system.zeta * r), v)
    return v


def slater_numpy(system, x, spin):
    n = system.n_up if spin == 0 else system.n_dn
    first = 0 if spin == 0 else system.n_up
    A = np.zeros((n, n))
    for c in range(n):
        xe = x[3 * (first + c):3 * (first + c) + 3]
        A[c] = system.coef[:n] @ basis_numpy(system, xe)
    return A


@pytest.mark.parametrize("Z", [2, 3, 5, 8, 10])
def test_lu_determinant_matches_leibniz(Z):
    system = S.atom(Z)
    wv = SJ.SjWave(system)
    L = SJ.lib()
    for x in cfgs(wv.N, 20, Z):
        for spin in (0, 1):
            sg = np.zeros(1)
            ld = L.orc_sj_logdet(wv.ref(), OF.dptr(np.ascontiguousarray(x)), spin, OF.dptr(sg))
            det = sg[0] * np.exp(ld)
            lz = float(L.orc_sj_det_leibniz(wv.ref(), OF.dptr(np.ascontiguousarray(x)), spin))
            # error scale: Hadamard's bound prod_c |row c|
            hb = np.prod(np.linalg.norm(slater_numpy(system, x, spin), axis=1)) if wv.N > 1 else 1.0
            assert abs(det - lz) <= 1e-13 * hb + 1e-300
            assert np.sign(det) == np.sign(lz) or abs(lz) < 1e-12 * hb


@pytest.mark.parametrize("Z", [3, 6, 10])
def test_inverse_times_matrix_is_identity(Z):
    wv = SJ.SjWave(S.atom(Z))
    sysd = S.atom(Z)
    for x in cfgs(wv.N, 10, 100 + Z):
        for spin in (0, 1):
            B, rc = SJ.inverse(wv, x, spin)
            assert rc == 0
            n = B.shape[0]
            A = slater_numpy(sysd, x, spin)
            assert np.allclose(B @ A, np.eye(n), atol=1e-9 * np.linalg.cond(A))


@pytest.mark.parametrize("system", [S.atom(3), S.atom(6), S.hydrogen_chain(4),
                                    S.random_molecule(7, seed=3)])
def test_local_energy_matches_finite_differences(system):
    wv = SJ.SjWave(system)
    h = 1e-4
    N = wv.N
    for x in cfgs(N, 6, 7):
        f0 = SJ.value(wv, x)
        lap = 0.0
        for k in range(3 * N):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            lap += (SJ.value(wv, xp) - 2 * f0 + SJ.value(wv, xm)) / h**2
        # potential
        R, Zs = system.nuclei, system.charges
        X = x.reshape(N, 3)
        V = -sum(Zs[I] / np.linalg.norm(X[i] - R[I]) for i in range(N) for I in range(len(Zs)))
        V += sum(1 / np.linalg.norm(X[i] - X[j]) for i in range(N) for j in range(i + 1, N))
        V += sum(Zs[I] * Zs[J] / np.linalg.norm(R[I] - R[J]) for I in range(len(Zs))
                 for J in range(I + 1, len(Zs)))
        want = -0.5 * lap / f0 + V
        got, _ = SJ.local_energy(wv, x)
        assert abs(got - want) <= 2e-5 * max(1.0, abs(want)), (got, want)


def test_zero_variance_eigenfunctions():
    # hydrogen-like: exp(-Z r) with Z = 1 and 3 gives E_L = -Z^2/2 everywhere
    for Z in (1, 3):
        wv = SJ.SjWave(S.hydrogenic(Z))
        for x in cfgs(1, 50, Z):
            e, _ = SJ.local_energy(wv, x)
            assert abs(e + Z * Z / 2) <= 1e-12
    # 2p: x_q exp(-Z r/2) gives E_L = -Z^2/8 everywhere (each axis); a p
    # laplacian with 2 z/r in place of 4 z/r would leave E_L = -Z^2/8 - Z/(2r)
    for Z in (1, 4):
        for axis in range(3):
            wv = SJ.SjWave(S.hydrogenic_2p(Z, axis))
            for x in cfgs(1, 30, 10 * Z + axis):
                e, _ = SJ.local_energy(wv, x)
                assert abs(e + Z * Z / 8) <= 1e-11 * max(1, Z * Z)
    # He, zeta = Z = 2, no Jastrow, IonicHamiltonian (operator.rs:130-149): E_L == -4
    wv = SJ.SjWave(S.helium_product(2.0, 0.0), hamiltonian=1)
    for x in cfgs(2, 50, 9):
        e, _ = SJ.local_energy(wv, x)
        assert abs(e + 4.0) <= 1e-12


def test_jastrow_parameter_log_derivatives():
    system = S.random_molecule(6, seed=5)
    h = 1e-6
    for x in cfgs(system.n_electrons, 5, 11):
        _, dl = SJ.local_energy(SJ.SjWave(system), x)
        for k in range(5):
            pp = np.array(system.jastrow, dtype=float)
            pm = pp.copy()
            pp[k] += h
            pm[k] -= h
            vp = SJ.value(SJ.SjWave(system, params=pp), x)
            vm = SJ.value(SJ.SjWave(system, params=pm), x)
            fd = (np.log(abs(vp)) - np.log(abs(vm))) / (2 * h)
            assert abs(dl[k] - fd) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("make", [lambda: S.random_molecule(6, seed=5), lambda: S.atom(8),
                                  lambda: S.harmonic_oscillator(1.0, a=1.0)])
def test_exponent_scale_log_derivative(make):
    # d ln psi / d s with the basis exponents s zeta_m: central differences of
    # ln|psi| over s (s-, p-type STO and GTO bases); the five Jastrow
    # derivatives are unchanged by the sixth parameter
    system = make()
    h = 1e-6
    for s0 in (1.0, 1.3):
        p6 = list(system.jastrow) + [s0]
        for x in cfgs(system.n_electrons, 4, 13, scale=0.9):
            e6, dl6 = SJ.local_energy(SJ.SjWave(system, params=p6), x)
            vp = SJ.value(SJ.SjWave(system, params=p6[:5] + [s0 + h]), x)
            vm = SJ.value(SJ.SjWave(system, params=p6[:5] + [s0 - h]), x)
            fd = (np.log(abs(vp)) - np.log(abs(vm))) / (2 * h)
            assert abs(dl6[5] - fd) <= 1e-6 * max(1.0, abs(fd)), (dl6[5], fd)
            if s0 == 1.0:
                e5, dl5 = SJ.local_energy(SJ.SjWave(system), x)
                assert e6 == e5 and np.array_equal(dl6[:5], dl5[:5]) and dl5[5] == 0.0


def test_exponent_scale_gaussian_closed_form():
    # psi = exp(-s zeta0 r^2): d ln psi / d s = -zeta0 r^2 exactly, and under
    # T + r^2/2 the local energy is 3 s zeta0 + r^2 (1/2 - 2 (s zeta0)^2)
    system = S.harmonic_oscillator(1.0, a=1.0)  # zeta0 = 1
    h = M_harmonic()
    rng = np.random.default_rng(3)
    for s0 in (0.3, 0.5, 1.0):
        wv = SJ.SjWave(system, hamiltonian=0, params=list(system.jastrow) + [s0])
        for _ in range(5):
            x = rng.normal(size=3)
            r2 = float(x @ x)
            _, dl = SJ.local_energy(wv, x)
            assert dl[5] == pytest.approx(-r2, rel=1e-14, abs=1e-15)
            e = SJ.operator_value(wv, x, h)
            z = s0
            assert e == pytest.approx(3 * z + r2 * (0.5 - 2 * z * z), rel=1e-12, abs=1e-12)


def M_harmonic():
    return [(SJ.TERM_KINETIC, 1.0), (SJ.TERM_HARMONIC, 1.0, 1.0, 0.0, 0.0, 0.0)]


@pytest.mark.parametrize("Z", [3, 7, 10])
def test_sherman_morrison_runner_follows_faithful_runner(Z):
    # same streams: SM ratios/updates (FP64, LU refresh every 16 moves) give the
    # same accept decisions as two full LU evaluations per move
    wv = SJ.SjWave(S.atom(Z))
    kw = dict(seed=2, walkers=24, steps=40, block_size=10, box_side=0.6)
    a = SJ.run(wv, faithful=True, **kw)
    b = SJ.run(wv, faithful=False, precision=SJ.PREC_FP64, lu_period=16, **kw)
    assert np.array_equal(a["accept_mask"], b["accept_mask"])
    assert np.array_equal(a["cfg"], b["cfg"])
    assert np.max(np.abs(a["e_local"] - b["e_local"]) / np.maximum(1, np.abs(a["e_local"]))) < 1e-8
    assert 0.2 < b["n_accepted"] / b["n_moves"] < 0.95


@pytest.mark.parametrize("Z", [3, 6])
def test_mixed_precision_runner_stays_close(Z):
    # FP32 Sherman-Morrison on well-conditioned determinants (Li, C: cond(A)
    # ~1e2..1e4) follows the FP64 chain; a walker whose decision flips diverges
    wv = SJ.SjWave(S.atom(Z))
    kw = dict(seed=3, walkers=32, steps=40, block_size=10, box_side=0.6, lu_period=64)
    a = SJ.run(wv, precision=SJ.PREC_FP64, **kw)
    b = SJ.run(wv, precision=SJ.PREC_MIXED, **kw)
    same = np.all(a["accept_mask"] == b["accept_mask"], axis=1)
    assert same.sum() >= 30
    rel = np.abs(a["e_local"][same] - b["e_local"][same]) / np.maximum(1, np.abs(a["e_local"][same]))
    assert np.max(rel) < 1e-3


def test_helium_product_energy_anchor_through_determinants():
    # tests/helium_lcao.rs:134: <E>(zeta) = zeta^2 - 27 zeta / 8
    zeta = 1.5
    wv = SJ.SjWave(S.helium_product(zeta, 0.0))
    r = SJ.run(wv, seed=4, walkers=256, steps=1200, block_size=100, box_side=1.0, lu_period=32,
               record=False, n_threads=8)
    mean = r["sum_e"] / r["n_samples"]
    nb = r["n_blocks"]
    err = np.sqrt((r["sum_bm2"] / nb - (r["sum_bm"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (zeta**2 - 27 * zeta / 8)) < 4 * err


@pytest.mark.parametrize("system", [S.atom(4), S.random_molecule(7, seed=3), S.helium_product(1.7, 0.5)])
def test_drift_matches_finite_differences_of_log_psi(system):
    # metrop.rs:153-155: drift = grad_i psi / psi
    wv = SJ.SjWave(system)
    h = 1e-5
    for x in cfgs(wv.N, 4, 13):
        for i in range(wv.N):
            want = np.zeros(3)
            for q in range(3):
                xp, xm = x.copy(), x.copy()
                xp[3 * i + q] += h
                xm[3 * i + q] -= h
                want[q] = (np.log(abs(SJ.value(wv, xp))) - np.log(abs(SJ.value(wv, xm)))) / (2 * h)
            got = SJ.drift(wv, x, i)
            assert np.allclose(got, want, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(want).max())), (i, got, want)


def test_diffusion_proposal_samples_psi_squared():
    # He product STO with the drift-diffusion proposal: the exact mean
    # zeta^2 - 27 zeta/8 (tests/helium_lcao.rs:134) within 4 sigma
    zeta = 1.5
    wv = SJ.SjWave(S.helium_product(zeta, 0.0))
    r = SJ.run(wv, seed=5, walkers=256, steps=1200, block_size=100, lu_period=32, record=False,
               n_threads=8, proposal=SJ.DIFFUSE, tau=0.3)
    mean = r["sum_e"] / r["n_samples"]
    nb = r["n_blocks"]
    err = np.sqrt((r["sum_bm2"] / nb - (r["sum_bm"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (zeta**2 - 27 * zeta / 8)) < 4 * err
    acc = r["n_accepted"] / r["n_moves"]
    assert 0.5 < acc < 1.0


def test_diffusion_acceptance_tends_to_one_for_small_steps():
    # the Green's-function ratio cancels psi'^2/psi^2 to O(tau^(3/2)) for a
    # nodeless psi (near a node the unlimited drift of metrop.rs:155 can
    # overshoot, as in the reference)
    wv = SJ.SjWave(S.helium_product(1.7, 0.5))
    accs = []
    for tau in (1e-2, 1e-4):
        r = SJ.run(wv, seed=6, walkers=64, steps=40, block_size=20, lu_period=16, record=False,
                   n_threads=8, proposal=SJ.DIFFUSE, tau=tau)
        accs.append(r["n_accepted"] / r["n_moves"])
    assert accs[1] > 0.999 and accs[1] > accs[0]


def test_diffusion_mixed_precision_runner_stays_close():
    wv = SJ.SjWave(S.atom(5))
    kw = dict(seed=2, walkers=16, steps=20, block_size=10, lu_period=16, n_threads=4,
              proposal=SJ.DIFFUSE, tau=0.05)
    a = SJ.run(wv, precision=SJ.PREC_FP64, **kw)
    b = SJ.run(wv, precision=SJ.PREC_MIXED, **kw)
    # FP32 Sherman-Morrison and Jastrow flip only rare near-threshold decisions
    assert np.mean(a["accept_mask"] == b["accept_mask"]) > 0.95


def test_basis_values_match_definitions():
    # the oracle's matrix entries against numpy's evaluation of every kind
    mixed = S.SJSystem([(0.0, 0.0, 0.0), (0.3, -0.2, 1.1)], [3, 1], 2, 1, [1.3, 0.7, 2.1, 2.1, 2.1, 0.9],
                       [0, 0, 0, 0, 0, 1], np.random.default_rng(0).normal(size=(2, 6)),
                       kind=[S.STO, S.GTO, S.STO_PX, S.STO_PY, S.STO_PZ, S.STO])
    wv = SJ.SjWave(mixed)
    for x in cfgs(3, 10, 21):
        for spin in (0, 1):
            B, rc = SJ.inverse(wv, x, spin)
            A = slater_numpy(mixed, x, spin)
            assert rc == 0
            assert np.allclose(B @ A, np.eye(A.shape[0]), atol=1e-10 * np.linalg.cond(A))


def wv_sys(Z):
    return S.atom(Z)


@pytest.mark.parametrize("Z", [3, 5, 8, 10])
def test_pivot_determinant_matches_lu_and_leibniz(Z):
    # the re-inversion's determinant (mant 2^ex from the Gauss-Jordan pivots)
    # against the independent LU and the long-double Leibniz expansion
    wv = SJ.SjWave(S.atom(Z))
    for x in cfgs(wv.N, 10, 40 + Z):
        for spin in (0, 1):
            n = wv.n_up if spin == 0 else wv.n_dn
            if n == 0:
                continue
            _, mant, ex, rc = SJ.inverse_det(wv, x, spin)
            assert rc == 0 and 0.5 <= abs(mant) < 1.0
            ld = SJ.logdet_value(mant, ex)
            lu, sg = SJ.logdet(wv, x, spin)
            # two backward-stable eliminations: |d log det| <~ n eps cond(A)
            cond = np.linalg.cond(slater_numpy(wv_sys(Z), x, spin))
            assert abs(ld - lu) <= 1e-15 * n * cond + 1e-13 * max(1.0, abs(lu))
            assert np.sign(mant) == sg
            lz = float(SJ.lib().orc_sj_det_leibniz(wv.ref(), OF.dptr(np.ascontiguousarray(x)), spin))
            # error scale of the expansion: Hadamard's bound prod_c |row c|
            hb = np.prod(np.linalg.norm(slater_numpy(wv_sys(Z), x, spin), axis=1))
            assert abs(mant * 2.0 ** ex - lz) <= 1e-13 * hb


def test_carried_logdet_of_parity_runner():
    # FP64: the determinant carried through the accepted moves since the last
    # re-inversion equals the LU determinant of the final configuration; the
    # drift recorded at re-inversions is at rounding level.  Mixed: the FP32
    # ratios leave a drift far above FP64's (what the device reports).
    wv = SJ.SjWave(S.atom(9))
    kw = dict(seed=5, walkers=16, steps=60, block_size=20, box_side=0.6, lu_period=64, record=False)
    a = SJ.run(wv, precision=SJ.PREC_FP64, **kw)
    b = SJ.run(wv, precision=SJ.PREC_MIXED, **kw)
    for w in range(16):
        for spin in (0, 1):
            lu, sg = SJ.logdet(wv, a["cfg"][w], spin)
            mant, ex = a["ld"][w, 2 * spin], a["ld"][w, 2 * spin + 1]
            assert abs(SJ.logdet_value(mant, ex) - lu) < 1e-10 and np.sign(mant) == sg
    assert 0.0 < a["max_ld_drift"] < 1e-10
    assert b["max_ld_drift"] > 1e3 * a["max_ld_drift"]
    assert a["n_singular"] == 0 and b["n_singular"] == 0


def test_singular_block_is_flagged_not_divided():
    # two spin-up electrons at the same point: identical rows, det A_up = 0.
    # The re-inversion flags the block (rc -1, mantissa 0) and its inverse
    # stays finite; the spin-down block is unaffected.
    wv = SJ.SjWave(S.atom(6))
    x = cfgs(6, 1, 3)[0]
    x[3:6] = x[0:3]
    B, mant, ex, rc = SJ.inverse_det(wv, x, 0)
    assert rc == -1 and mant == 0.0 and np.all(np.isfinite(B))
    assert SJ.logdet_value(mant, ex) == -np.inf
    _, mant_d, _, rc_d = SJ.inverse_det(wv, x, 1)
    assert rc_d == 0 and mant_d != 0.0
    # s functions only: two electrons at the same distance from the nucleus
    # are identical rows as well
    s_only = S.SJSystem([(0.0, 0.0, 0.0)], [4], 2, 2, [1.1, 3.0], [0, 0],
                        np.random.default_rng(1).normal(size=(2, 2)))
    wv2 = SJ.SjWave(s_only)
    y = cfgs(4, 1, 4)[0]
    y[3:6] = y[0:3][[1, 2, 0]]  # a rotation: same |r|
    _, m2, _, rc2 = SJ.inverse_det(wv2, y, 0)
    assert rc2 == -1 and m2 == 0.0
    # the runner counts such blocks at its re-inversions (here the first one)
    cfg = np.tile(x, (4, 1))
    r = SJ.run(wv, seed=1, walkers=4, steps=4, block_size=2, box_side=0.5, lu_period=64, cfg=cfg, record=False)
    assert r["n_singular"] == 4


# exact <E> of the H2+ product function at R = 2.5 and its local-energy
# standard deviation, by quadrature in cylindrical coordinates (scipy dblquad,
# 1e-11 relative): the anchor of tests/hydrogen_molecular_ion_lcao.rs:100-142
H2ION_E = -0.43720930232581723
H2ION_STD = 0.5052200457935744


def test_h2_ion_product_is_the_reference_function():
    # psi and E_L of the Slater-Jastrow form against the reference's H2WF
    # (tests/hydrogen_molecular_ion_lcao.rs:60-98) written out in numpy
    wv = SJ.SjWave(S.h2_ion_product())
    A, B = np.array([-1.25, 0, 0]), np.array([1.25, 0, 0])
    for x in cfgs(1, 20, 31, scale=1.5):
        r1, r2 = np.linalg.norm(x - A), np.linalg.norm(x - B)
        assert abs(SJ.value(wv, x) - np.exp(-r1) * np.exp(-r2)) <= 1e-15 * np.exp(-r1 - r2) * 4
        g = -(x - A) / r1 - (x - B) / r2
        want = -0.5 * (g @ g - 2 / r1 - 2 / r2) - 1 / r1 - 1 / r2 + 1 / 2.5
        e, _ = SJ.local_energy(wv, x)
        assert abs(e - want) <= 1e-12 * max(1, abs(want))


def test_h2_ion_product_energy_anchor():
    # the reference asserts |E + 0.565| < std of the samples; the oracle's mean
    # also sits on the quadrature value within its blocking error
    wv = SJ.SjWave(S.h2_ion_product())
    r = SJ.run(wv, seed=1, walkers=256, steps=1000, block_size=100, box_side=1.0, lu_period=64, record=False,
               n_threads=8)
    n = r["n_samples"]
    mean = r["sum_e"] / n
    std = np.sqrt(r["sum_e2"] / n - mean * mean)
    nb = r["n_blocks"]
    err = np.sqrt((r["sum_bm2"] / nb - (r["sum_bm"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (-0.565)) < std
    assert abs(mean - H2ION_E) < 4 * err
    assert abs(std - H2ION_STD) < 0.05 * H2ION_STD


@pytest.mark.parametrize("params", [None, (0.5, 0.25, 1.0, 0.3, 1.0, 1.07)])
def test_compat_psi_samples_are_the_wavefunction_value(params):
    # the parity runner's "Wavefunction value" samples, (psi psi)/psi from the
    # carried determinants and J = b_anti O_0 + b_para O_1 + a O_3: the last
    # sample of a walker is taken at its final configuration, where the LU
    # determinants and the full Jastrow sum of orc_sj_value give psi afresh
    system = S.random_molecule(6, seed=5)
    wv = SJ.SjWave(system) if params is None else SJ.SjWave(system, params=np.array(params))
    r = SJ.run(wv, seed=3, walkers=12, steps=40, block_size=10, box_side=0.6, lu_period=64,
               precision=SJ.PREC_FP64)
    assert r["psi"].shape == (12, 30) and np.all(np.isfinite(r["psi"]))
    for w in range(12):
        v = SJ.value(wv, r["cfg"][w])
        assert np.sign(r["psi"][w, -1]) == np.sign(v)
        assert abs(r["psi"][w, -1] - v) <= 1e-9 * abs(v)
