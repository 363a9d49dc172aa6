"""Pins the oracle of the learned MH proposal: the input features against a
brute-force restatement, the network against numpy, the analytic training
gradient against finite differences, the proposal densities against the
reversibility identity of Metropolis-Hastings, and the determinant-ratio and
drift pieces against the wavefunction value."""
import numpy as np
import pytest

from mole_amd import systems as S
from oracle import mlp_ffi as ML
from oracle import sj_ffi as SJ

from helpers_learned import exact_hydrogen_configs, hydrogen_radial_chi2, shaped_weights


@pytest.fixture(scope="module")
def ne():
    return SJ.SjWave(S.atom(10))


def cfgs(N, n, seed, scale=1.0):
    return np.random.default_rng(seed).normal(scale=scale, size=(n, 3 * N))


def test_features_brute_force(ne):
    for x in cfgs(10, 5, 1):
        X = x.reshape(10, 3)
        for i in range(10):
            f = ML.features(ne, x, i)
            assert f[0] == pytest.approx(np.linalg.norm(X[i]), rel=1e-15)
            d = np.linalg.norm(X - X[i], axis=1)
            d[i] = np.inf
            order = np.argsort(d, kind="stable")[:3]
            want = np.concatenate([X[j] - X[i] for j in order])
            assert np.allclose(f[1:10], want, rtol=0, atol=1e-15)
            assert np.array_equal(f[10:], X[i])


def test_features_pad_missing_partners():
    he = SJ.SjWave(S.helium_product(1.7, 0.5))
    x = np.array([0.1, 0.2, 0.3, -0.4, 0.5, 0.6])
    f = ML.features(he, x, 0)
    assert np.allclose(f[1:4], [-0.5, 0.3, 0.3]) and np.all(f[4:10] == 0.0)


def test_forward_matches_numpy():
    w = ML.xavier(2)
    rng = np.random.default_rng(3)
    f = rng.normal(size=13)
    W1 = w[0:832].reshape(13, 64).astype(float); b1 = w[832:896]
    W2 = w[896:4992].reshape(64, 64).astype(float); b2 = w[4992:5056]
    W3 = w[5056:9152].reshape(64, 64).astype(float); b3 = w[9152:9216]
    W4 = w[9216:9472].reshape(64, 4).astype(float); b4 = w[9472:9476]
    w[832:896] = rng.normal(size=64) * 0.1
    b1 = w[832:896].astype(float)
    h = np.tanh(f @ W1 + b1)
    h = np.tanh(h @ W2 + b2)
    h = np.tanh(h @ W3 + b3)
    out = h @ W4 + b4
    got, act = ML.forward(w, f)
    assert np.allclose(got, out, rtol=1e-12, atol=1e-13)
    # bf16 emulation stays within bf16 rounding of the FP64 network
    got16, _ = ML.forward(w, f, emulate_bf16=True)
    assert np.max(np.abs(got16 - out)) < 0.05


def test_log_sigma_is_a_smooth_bound():
    # ln sigma = ln 0.25 + 3 tanh(out / 3): the textbook tanh, the bounds, and
    # a slope that stays positive far outside the old hard clamp [-7, 1]
    for o in [-40.0, -9.0, -7.5, -1.0, 0.0, 0.4, 1.0, 1.5, 4.0, 60.0]:
        assert ML.log_sigma(o) == pytest.approx(ML.LS0 + 3.0 * np.tanh(o / 3.0), abs=2e-15)
        h = 1e-6
        fd = (ML.log_sigma(o + h) - ML.log_sigma(o - h)) / (2 * h)
        assert ML.log_sigma_slope(o) == pytest.approx(fd, rel=1e-6, abs=1e-9)
    assert ML.log_sigma(0.0) == ML.LS0 and ML.log_sigma_slope(0.0) == 1.0
    assert np.exp(ML.log_sigma(1e6)) == pytest.approx(0.25 * np.exp(3.0))
    assert np.exp(ML.log_sigma(-1e6)) == pytest.approx(0.25 * np.exp(-3.0))
    assert ML.log_sigma_slope(1.5) > 0.7 and ML.log_sigma_slope(-9.0) > 0.009


@pytest.mark.parametrize("b43,objective", [(0.3, ML.OBJ_ESJD), (2.5, ML.OBJ_ESJD),
                                           (0.3, ML.OBJ_ESJD_RELATIVE), (0.3, ML.OBJ_LOG_ESJD)])
def test_training_gradient_matches_finite_differences(ne, b43, objective):
    # b4[3] = 2.5 puts ln sigma's output where a hard clamp would give no slope
    w = ML.xavier(2)
    rng = np.random.default_rng(4)
    w[ML.OFF["b4"] + 3] = b43
    samples = []
    for x in cfgs(10, 12, 5, scale=0.8):
        i = int(rng.integers(0, 10))
        z = rng.normal(size=3)
        O, g, d = ML.sample(ne, w, x, i, z, grad=True, objective=objective)
        if abs(d["logr"]) < 0.05:  # keep away from the kink of min(1, r)
            continue
        cfg2 = x.copy()
        cfg2[3 * i:3 * i + 3] = d["xn"]
        samples.append((x, i, z, ML.features(ne, cfg2, i, d["xn"]), g))
    assert len(samples) >= 6
    w64 = w.astype(np.float64)
    idx = [0, 17, 831, 850, 900, 3000, 5000, 7000, 9152, 9220, 9400, 9472, 9475]
    for x, i, z, f2, g in samples:
        for p in idx:
            h = 1e-3 * max(1.0, abs(w64[p]))
            wp, wm = w64.copy(), w64.copy()
            wp[p] += h
            wm[p] -= h
            # float32 weights: the perturbation itself is rounded, use the rounded step
            wpf, wmf = wp.astype(np.float32), wm.astype(np.float32)
            hh = (float(wpf[p]) - float(wmf[p])) / 2
            Op, _, _ = ML.sample(ne, wpf, x, i, z, frozen_f2=f2, objective=objective)
            Om, _, _ = ML.sample(ne, wmf, x, i, z, frozen_f2=f2, objective=objective)
            fd = (Op - Om) / (2 * hh)
            assert abs(fd - g[p]) <= 2e-4 * max(1.0, abs(fd)) + 1e-6, (p, fd, g[p])


def test_objectives_relate_as_documented(ne):
    # one sample under the three objectives: RELATIVE is ESJD over
    # d^2 + 0.01 (d the larger float-rounded nearest-nucleus distance of x
    # and x'), LOG is log(1e-4 + ESJD); the gradients scale the same way
    w = ML.xavier(2)
    rng = np.random.default_rng(21)
    for x in cfgs(10, 4, 22, scale=0.8):
        i = int(rng.integers(0, 10))
        z = rng.normal(size=3)
        O0, g0, d = ML.sample(ne, w, x, i, z, grad=True)
        assert O0 == pytest.approx(d["A"] * d["d2"], rel=1e-15)
        cfg2 = x.copy()
        cfg2[3 * i:3 * i + 3] = d["xn"]
        da = float(np.float32(np.linalg.norm(x[3 * i:3 * i + 3])))
        db = float(np.float32(np.linalg.norm(d["xn"])))
        wgt = 1.0 / (max(da, db) ** 2 + 0.01)
        O1, g1, _ = ML.sample(ne, w, x, i, z, grad=True, objective=ML.OBJ_ESJD_RELATIVE)
        assert O1 == pytest.approx(O0 * wgt, rel=1e-12)
        assert np.allclose(g1, g0 * wgt, rtol=1e-9, atol=1e-14)
        O2, g2, _ = ML.sample(ne, w, x, i, z, grad=True, objective=ML.OBJ_LOG_ESJD)
        assert O2 == pytest.approx(np.log(1e-4 + O0), rel=1e-14)
        assert np.allclose(g2, g0 / (1e-4 + O0), rtol=1e-9, atol=1e-12)


def test_mh_reversibility_identity(ne):
    # psi(x)^2 q(x'|x) A(x->x') = psi(x')^2 q(x|x') A(x'->x) for the ratio the
    # sampler uses: log q(x|x') - log q(x'|x) is antisymmetric under x <-> x'
    w = ML.xavier(7)
    rng = np.random.default_rng(8)
    for x in cfgs(10, 6, 9, scale=0.8):
        i = int(rng.integers(0, 10))
        xn = x[3 * i:3 * i + 3] + rng.normal(scale=0.3, size=3)
        x2 = x.copy()
        x2[3 * i:3 * i + 3] = xn
        a = ML.log_q_ratio(ne, w, x, i, xn)
        b = ML.log_q_ratio(ne, w, x2, i, x[3 * i:3 * i + 3])
        assert abs(a + b) <= 1e-12 * max(1.0, abs(a))
        lR, _ = ML.move_terms(ne, x, i, xn)
        lR_back, _ = ML.move_terms(ne, x2, i, x[3 * i:3 * i + 3])
        # the synthetic Ne determinants have condition numbers near 1e10
        assert abs(lR + lR_back) <= 1e-6 * max(1.0, abs(lR))
        # both directions of the balance equation
        lhs = np.log(SJ.value(ne, x) ** 2) + min(0.0, lR + a) - a / 2
        rhs = np.log(SJ.value(ne, x2) ** 2) + min(0.0, -lR - a) + a / 2
        # with log q(x'|x) = -a/2 + c and log q(x|x') = a/2 + c
        assert abs(lhs - rhs) <= 1e-6 * max(1.0, abs(lhs))


@pytest.mark.parametrize("Z,tol", [(3, 1e-11), (4, 1e-8)])
def test_move_terms_against_values_and_drift(Z, tol):
    # tolerance follows the conditioning of the synthetic determinants
    ne = SJ.SjWave(S.atom(Z))
    rng = np.random.default_rng(10)
    for x in cfgs(Z, 4, 11, scale=0.8):
        i = int(rng.integers(0, Z))
        xn = x[3 * i:3 * i + 3] + rng.normal(scale=0.2, size=3)
        x2 = x.copy()
        x2[3 * i:3 * i + 3] = xn
        lR2, g = ML.move_terms(ne, x, i, xn)
        assert lR2 == pytest.approx(np.log((SJ.value(ne, x2) / SJ.value(ne, x)) ** 2), rel=tol, abs=tol)
        assert np.allclose(g, SJ.drift(ne, x2, i), rtol=tol, atol=tol)


def test_learned_sampler_keeps_exact_distribution_and_a_planted_mistake_does_not():
    # stationarity on a one-electron target: start from exact psi^2 draws
    h = SJ.SjWave(S.hydrogenic(1))
    w = shaped_weights()
    W = 8192
    cfg0 = exact_hydrogen_configs(W, 1)
    lim = None
    for drop in (False, True):
        r = SJ.run(h, seed=4, walkers=W, steps=200, block_size=10, lu_period=16, n_threads=8, record=False,
                   proposal=SJ.LEARNED, mlp_w=w, cfg=cfg0, pos0=0, drop_q_ratio=drop)
        chi2, dof = hydrogen_radial_chi2(np.linalg.norm(r["cfg"].reshape(-1, 3), axis=1))
        lim = dof + 6 * np.sqrt(2 * dof)
        if not drop:
            assert chi2 < lim, chi2
            assert 0.3 < r["n_accepted"] / r["n_moves"] < 0.95
        else:
            assert chi2 > 20 * lim, chi2  # omitting q(x|x')/q(x'|x) is caught
