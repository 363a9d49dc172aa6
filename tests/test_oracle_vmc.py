"""Pins the oracle's Sampler/Runner/VmcRunner restatement and reductions:
invariants of the box Metropolis step, the reference's statistical anchors,
closed forms of blocking and tau, and brute-force SR/gradient sums."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ffi as O


def test_uniform_psi_always_accepts():
    # metrop.rs:259-268: a constant psi accepts every proposal
    wf = O.helium(0.0, 0.0, 0.0)
    r = O.run_box(wf, seed=7, walkers=8, steps=40, block_size=10, box_side=1.0)
    assert r["n_accepted"] == r["n_moves"] == 8 * 40 * 2
    assert np.all(r["accept_mask"] == 3)
    assert np.allclose(r["acceptance"], 40.0)  # 1/N per accepted move


def test_runner_asserts_steps_at_least_two_blocks():
    with pytest.raises(ValueError):
        O.run_box(O.helium(), seed=0, walkers=1, steps=19, block_size=10, box_side=0.5)


def test_sample_count_and_block_discard():
    r = O.run_box(O.helium(), seed=1, walkers=3, steps=57, block_size=10, box_side=0.5)
    # blocks = 5; the first block is equilibration (montecarlo.rs:36)
    assert r["e_local"].shape == (3, 40)
    assert r["n_samples"] == 120 and r["n_blocks"] == 12
    assert r["n_moves"] == 3 * 50 * 2


def test_deterministic_and_thread_count_independent():
    a = O.run_box(O.h2(), seed=5, walkers=16, steps=60, block_size=20, box_side=0.5, n_threads=1)
    b = O.run_box(O.h2(), seed=5, walkers=16, steps=60, block_size=20, box_side=0.5, n_threads=4)
    assert np.array_equal(a["accept_mask"], b["accept_mask"])
    assert np.array_equal(a["e_local"], b["e_local"])
    assert np.array_equal(a["cfg"], b["cfg"])


def test_continuation_equals_single_run():
    # (cfg, stream position) determine the continuation: k + (N-k) == N sweeps
    wf = O.helium()
    full = O.run_box(wf, seed=9, walkers=4, steps=40, block_size=20, box_side=0.5)
    part = O.run_box(wf, seed=9, walkers=4, steps=20, block_size=10, box_side=0.5)
    rest = O.run_box(wf, seed=9, walkers=4, steps=20, block_size=10, box_side=0.5,
                     cfg=part["cfg"], pos0=6 + 20 * 2 * 4)
    assert np.array_equal(full["cfg"], rest["cfg"])
    assert np.array_equal(full["accept_mask"][:, 20:], rest["accept_mask"])


@pytest.mark.parametrize("zeta", [1.5, 27.0 / 16.0, 2.0])
def test_helium_product_sto_energy_anchor(zeta):
    # tests/helium_lcao.rs:134: <E>(zeta) = zeta^2 - 27 zeta / 8 for b = 0
    wf = O.helium(zeta, 0.0, 0.3)
    W, S, B = 256, 2000, 100
    r = O.run_box(wf, seed=11, walkers=W, steps=S, block_size=B, box_side=1.0,
                  record=False, n_threads=8)
    mean = r["sum_e"] / r["n_samples"]
    nb = r["n_blocks"]
    err = np.sqrt((r["sum_bm2"] / nb - (r["sum_bm"] / nb) ** 2) / (nb - 1))
    exact = zeta**2 - 27.0 * zeta / 8.0
    assert abs(mean - exact) < 4.0 * err, (mean, exact, err)


def test_running_mean_and_blocking_closed_forms():
    L = O.lib()
    rng = np.random.default_rng(0)
    x = rng.normal(1.0, 2.0, 200000)
    assert abs(L.orc_running_mean(O.dptr(x), x.size) - x.mean()) <= 1e-12
    # iid data: blocking error ~ sigma / sqrt(n)
    err = L.orc_blocking_error(O.dptr(x), x.size, 100)
    assert abs(err / (2.0 / np.sqrt(x.size)) - 1.0) < 0.15
    # brute force of vmc.rs:150-170
    bm = x.reshape(-1, 100).mean(axis=1)
    want = np.sqrt(((bm**2).mean() - x.mean() ** 2) / (bm.size - 1))
    assert abs(err - want) <= 1e-10 * want


def test_tau_of_ar1_series():
    L = O.lib()
    rng = np.random.default_rng(1)
    rho = 0.8
    n = 400000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0]
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    tau = L.orc_tau_corr(O.dptr(x), n)
    assert abs(tau - (1 + rho) / (1 - rho)) < 0.5
    iid = rng.normal(size=20000)
    assert L.orc_tau_corr(O.dptr(iid), iid.size) < 1.3


def test_energy_gradient_and_sr_matrix_brute_force():
    L = O.lib()
    rng = np.random.default_rng(2)
    n, npar = 5000, 3
    G = rng.normal(size=(n, npar))
    psi = rng.uniform(0.5, 2.0, n)
    E = rng.normal(-1.0, 0.3, n)
    ebar = E.mean()
    g = np.zeros(npar)
    L.orc_energy_gradient(O.dptr(G), O.dptr(psi), O.dptr(E), n, npar, ebar, O.dptr(g))
    Ov = G / psi[:, None]
    assert np.allclose(g, 2 * (Ov * (E - ebar)[:, None]).mean(axis=0), rtol=1e-12, atol=1e-14)
    S = np.zeros(npar * npar)
    L.orc_sr_matrix(O.dptr(G), O.dptr(psi), n, npar, O.SR_TEXTBOOK, 1e-3, O.dptr(S))
    want = Ov.T @ Ov / n - np.outer(Ov.mean(0), Ov.mean(0))
    want[np.diag_indices(npar)] *= 1 + 1e-3
    assert np.allclose(S.reshape(npar, npar), want, rtol=1e-10, atol=1e-13)
    L.orc_sr_matrix(O.dptr(G), O.dptr(psi), n, npar, O.SR_REFERENCE, 0.0, O.dptr(S))
    want = Ov.T @ Ov / n - Ov.mean(0).sum() ** 2  # Q2: scalar on every element
    want[np.diag_indices(npar)] *= 1.01
    assert np.allclose(S.reshape(npar, npar), want, rtol=1e-10, atol=1e-13)
    # solve
    A = rng.normal(size=(npar, npar))
    A = A @ A.T + npar * np.eye(npar)
    b = rng.normal(size=npar)
    x = np.zeros(npar)
    assert L.orc_solve_sym(O.dptr(A.ravel().copy()), O.dptr(b), npar, O.dptr(x)) == 0
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12)


def test_vmc_sr_singular_matrix_is_linalg_error():
    # b = 0 leaves c without effect: S is singular, solveh_into fails
    wf = O.helium(1.4, 0.0, 0.3)
    op = O.OptParams(optimizer=O.OPT_SR, step=0.5, sr_mode=O.SR_TEXTBOOK, sr_eps=0.0,
                     o_convention=O.O_LOGDERIV, persistent=1)
    z = [np.zeros(3) for _ in range(5)]
    rc = O.lib().orc_vmc_optimize(C.byref(wf), 3, 1.0, 1, 64 * 100, 20, 64, C.byref(op), 4,
                                  *[O.dptr(a) for a in z])
    assert rc == -2


def test_vmc_sr_lowers_helium_energy():
    wf = O.helium(1.4, 0.5, 0.3)
    iters = 12
    op = O.OptParams(optimizer=O.OPT_SR, step=0.1, sr_mode=O.SR_TEXTBOOK, sr_eps=1e-3,
                     o_convention=O.O_LOGDERIV, persistent=1)
    en, er, ac, gr, pa = (np.zeros(iters), np.zeros(iters), np.zeros(iters),
                          np.zeros(iters * 3), np.zeros(iters * 3))
    rc = O.lib().orc_vmc_optimize(C.byref(wf), 3, 1.0, iters, 64 * 400, 40, 64, C.byref(op), 8,
                                  O.dptr(en), O.dptr(er), O.dptr(ac), O.dptr(gr), O.dptr(pa))
    assert rc == 0
    zeta = pa.reshape(iters, 3)[:, 0]
    assert zeta[-1] > 1.55
    assert en[-1] < en[0] - 5 * er[0]
    assert en[-1] < -2.86  # below the product-STO optimum -(27/16)^2
    assert np.all((ac > 0.3) & (ac < 1.0))


def test_vectorised_tau_matches_the_oracle_estimator():
    from mole_amd.stats import tau_int
    rng = np.random.default_rng(21)
    series = []
    for rho in (0.0, 0.5, 0.9, 0.97):
        for n in (50, 300, 1000):
            e = rng.normal(size=n)
            x = np.empty(n)
            x[0] = e[0]
            for t in range(1, n):
                x[t] = rho * x[t - 1] + e[t]
            series.append(x)
    for x in series:
        want = O.lib().orc_tau_corr(O.dptr(np.ascontiguousarray(x)), x.size)
        assert tau_int(x)[0] == pytest.approx(want, rel=1e-9, abs=1e-9)
    # batched rows equal single rows
    X = np.stack([s for s in series if s.size == 300])
    assert np.allclose(tau_int(X), [tau_int(r)[0] for r in X], rtol=1e-12)
