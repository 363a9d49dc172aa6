"""Pins the oracle's Slater-Jastrow restatement: LU determinants against the
Leibniz expansion, Gauss-Jordan inverses against A B = I, the local energy
against finite differences and zero-variance eigenfunctions, the Jastrow
parameter derivatives against finite differences, and the Sherman-Morrison
runner against the from-scratch (reference-faithful) runner."""
import ctypes as C

import numpy as np
import pytest

from mole_amd import systems as S
from oracle import oracle_ffi as OF
from oracle import sj_ffi as SJ


def cfgs(N, n, seed, scale=1.2):
    return np.random.default_rng(seed).normal(scale=scale, size=(n, 3 * N))


def slater_numpy(system, x, spin):
    n = system.n_up if spin == 0 else system.n_dn
    first = 0 if spin == 0 else system.n_up
    A = np.zeros((n, n))
    for c in range(n):
        xe = x[3 * (first + c):3 * (first + c) + 3]
        r = np.linalg.norm(xe[None, :] - system.nuclei[system.center], axis=1)
        A[c] = system.coef[:n] @ np.exp(-system.zeta * r)
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
