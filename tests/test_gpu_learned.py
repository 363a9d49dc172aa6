"""Learned Metropolis-Hastings proposal on the gfx950: the bf16 MFMA network
against the oracle's FP64 network (with and without bf16 operand rounding),
the device training gradient against the oracle's analytic gradient of the
same samples, exactness of the sampler on a one-electron target (histogram
against the exact radial density) and energy agreement with the box
proposal."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import systems as S
from oracle import mlp_ffi as ML
from oracle import oracle_ffi as OF
from oracle import sj_ffi as SJ

from helpers_learned import exact_hydrogen_configs, hydrogen_radial_chi2, shaped_weights

pytestmark = pytest.mark.gpu

TRAIN_SYS = 0x6D6C70


def test_device_xavier_init():
    with M.Mlp(seed=2) as m:
        w = m.weights()
    for name, (nin, nout) in dict(W1=(13, 64), W2=(64, 64), W3=(64, 64), W4=(64, 4)).items():
        a = np.sqrt(6.0 / (nin + nout))
        blk = w[ML.OFF[name]:ML.OFF[name] + nin * nout]
        assert np.all(np.abs(blk) <= a) and abs(blk.mean()) < 0.1 * a
        assert abs(blk.std() - a / np.sqrt(3)) < 0.15 * a
    for b in ("b1", "b2", "b3", "b4"):
        n = 64 if b != "b4" else 4
        assert np.all(w[ML.OFF[b]:ML.OFF[b] + n] == 0.0)
    with M.Mlp(seed=2) as m2, M.Mlp(seed=3) as m3:
        assert np.array_equal(m2.weights(), w) and not np.array_equal(m3.weights(), w)


def _random_weights(seed):
    w = ML.xavier(seed)
    rng = np.random.default_rng(seed)
    for b in ("b1", "b2", "b3"):
        w[ML.OFF[b]:ML.OFF[b] + 64] = rng.normal(scale=0.2, size=64)
    w[ML.OFF["b4"]:ML.OFF["b4"] + 4] = rng.normal(scale=0.3, size=4)
    return w


def test_forward_matches_oracle_network():
    wv = SJ.SjWave(S.atom(10))
    w = _random_weights(5)
    rng = np.random.default_rng(6)
    n = 200
    cfg = rng.normal(scale=0.9, size=(n, 30))
    el = rng.integers(0, 10, n).astype(np.int32)
    with M.Mlp() as m:
        m.set_weights(w)
        assert np.array_equal(m.weights(), w)
        got = m.eval(cfg, [0.0, 0.0, 0.0], el)
    want16 = np.array([ML.forward(w, ML.features(wv, cfg[e], el[e]), emulate_bf16=True)[0] for e in range(n)])
    want64 = np.array([ML.forward(w, ML.features(wv, cfg[e], el[e]))[0] for e in range(n)])
    # bf16 operands, FP32 accumulation and tanhf against bf16-rounded FP64
    assert np.max(np.abs(got - want16)) < 0.02
    assert np.max(np.abs(got - want64)) < 0.1
    assert np.std(got[:, 3]) > 0.05  # a non-trivial network


def _ne_context(W, seed=1):
    dev = [M.slater_jastrow(S.atom(10), W)]
    return M.Context(dev, seed=seed, step=0.6, block_size=2, precision=M.PREC_FP64, lu_period=16)


@pytest.mark.parametrize("objective", [M.MLP_OBJ_ESJD, M.MLP_OBJ_ESJD_RELATIVE, M.MLP_OBJ_LOG_ESJD])
def test_training_gradient_matches_oracle(objective):
    wv = SJ.SjWave(S.atom(10))
    w = _random_weights(9)
    W, seed, batch = 64, 17, 512
    with _ne_context(W) as ctx, M.Mlp() as m:
        m.set_weights(w)
        ctx.init_walkers()
        data = []
        for _ in range(2):
            m.collect(ctx, 1, 3)
            data.append(ctx.get_configs())
        data = np.concatenate(data).reshape(-1, 30)
        assert m.data_size() == data.shape[0]
        g_dev, o_dev = m.gradient(ctx, batch, first=0, seed=seed, objective=objective)
    L = OF.lib()
    g_ref = np.zeros(ML.NPARAM)
    o_ref = 0.0
    N = 10
    for g in range(batch):
        i, c = g % N, (g // N) % data.shape[0]
        d = [L.orc_stream_u64(seed, TRAIN_SYS, g & 0xFFFFFFFF, g >> 32, k) for k in range(4)]
        z0, z1 = OF.box_muller(d[0], d[1])
        z2, _ = OF.box_muller(d[2], d[3])
        O, gr, _ = ML.sample(wv, w, data[c], i, np.array([z0, z1, z2]), emulate_bf16=True, grad=True,
                             objective=objective)
        o_ref += O
        g_ref += gr
    assert abs(o_dev - o_ref) <= 0.03 * abs(o_ref)
    cos = g_dev @ g_ref / (np.linalg.norm(g_dev) * np.linalg.norm(g_ref))
    rel = np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref)
    assert cos > 0.99 and rel < 0.15, (cos, rel)


def test_learned_proposal_keeps_hydrogen_exact():
    # one electron, psi = exp(-r): start from exact psi^2 draws and check the
    # radial distribution stays exact under an asymmetric learned proposal
    # (the CPU test of the same setup shows a wrong ratio fails it); E_L == -1/2
    W = 8192
    dev = [M.slater_jastrow(S.hydrogenic(1), W)]
    with M.Context(dev, seed=4, step=0.0, block_size=10, proposal=M.PROPOSAL_LEARNED, lu_period=16,
                   flags=M.FLAG_TRACE, record_sweeps=200) as ctx, M.Mlp() as m:
        m.set_weights(shaped_weights())
        m.attach(ctx)
        ctx.set_configs(exact_hydrogen_configs(W, 1).ravel())
        ctx.run(200)
        st = ctx.stats()
        cfg = ctx.get_configs().reshape(W, 3)
        _, el = ctx.trace(200, 190)
    assert np.max(np.abs(el + 0.5)) <= 1e-12
    assert 0.3 < st["n_accepted"] / st["n_moves"] < 0.95
    chi2, dof = hydrogen_radial_chi2(np.linalg.norm(cfg, axis=1))
    assert chi2 < dof + 6 * np.sqrt(2 * dof), chi2


def test_learned_energy_agrees_with_box():
    # He product state, exact mean zeta^2 - 27 zeta / 8.  The asymmetric
    # shaped proposal (acceptance ~0.56, tau_int ~13 sweeps on the oracle)
    # starts from a box-equilibrated ensemble: an untrained network mixes
    # too slowly (acceptance ~0.09) for 1000 sweeps to leave its burn-in.
    system = S.helium_product(1.5, 0.0)
    W = 8192
    dev = [M.slater_jastrow(system, W)]
    with M.Context(dev, seed=5, step=1.0, block_size=100, lu_period=32) as box:
        box.init_walkers()
        box.sweep(400)
        start = box.get_configs()
    with M.Context(dev, seed=6, step=0.0, block_size=100, proposal=M.PROPOSAL_LEARNED,
                   lu_period=32) as ctx, M.Mlp() as m:
        m.set_weights(shaped_weights())
        m.attach(ctx)
        ctx.set_configs(start)
        ctx.sweep(100)
        ctx.run(1000)
        st = ctx.stats()
    mean = st["sum_e"] / st["n_samples"]
    nb = st["n_blocks"]
    err = np.sqrt((st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2) / (nb - 1))
    zeta = 1.5
    assert abs(mean - (zeta**2 - 27 * zeta / 8)) < 4 * err


def test_training_raises_the_objective():
    W = 2048
    with _ne_context(W, seed=3) as ctx, M.Mlp(seed=2) as m:
        ctx.init_walkers()
        ctx.sweep(50)
        m.collect(ctx, 4, 5)
        obj = m.train(ctx, 150, batch=2048, lr=2e-3, seed=5)
    assert np.all(np.isfinite(obj))
    assert obj[-20:].mean() > 1.2 * obj[:10].mean(), (obj[:10].mean(), obj[-20:].mean())


def _adam_f32(w, g_sum, m, v, step, batch, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32, the operations of the device step: the loss
    is minus the batch mean of the objective, step counts from 1."""
    f = np.float32
    w, m, v = w.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
    gr = -g_sum.astype(np.float32) * (f(1.0) / f(batch))
    b1, b2, lr, eps = f(b1), f(b2), f(lr), f(eps)
    m = b1 * m + (f(1.0) - b1) * gr
    v = b2 * v + (f(1.0) - b2) * gr * gr
    mh = m / (f(1.0) - b1 ** f(step))
    vh = v / (f(1.0) - b2 ** f(step))
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def test_adam_steps_match_float32_restatement():
    # two device training steps against Adam restated in numpy float32 from
    # the device's own gradients of the same samples (step k uses samples
    # [k batch, (k+1) batch)): bias correction, the betas and the moments
    # carried between steps all enter; the repacked bf16 network follows
    W, batch, seed = 256, 512, 21
    rng = np.random.default_rng(3)
    cfg = rng.normal(scale=0.9, size=(32, 30))
    el = rng.integers(0, 10, 32).astype(np.int32)
    with _ne_context(W, seed=2) as ctx, M.Mlp(seed=2) as m:
        ctx.init_walkers()
        ctx.sweep(10)
        m.collect(ctx, 2, 3)
        w0 = _random_weights(4)
        m.set_weights(w0)
        out0 = m.eval(cfg, [0.0, 0.0, 0.0], el)
        g0, _ = m.gradient(ctx, batch, first=0, seed=seed)
        m.train(ctx, 1, batch=batch, lr=1e-3, seed=seed)
        w1_dev = m.weights()
        g1, _ = m.gradient(ctx, batch, first=batch, seed=seed)
        m.train(ctx, 1, batch=batch, lr=1e-3, seed=seed)
        w2_dev = m.weights()
        out2 = m.eval(cfg, [0.0, 0.0, 0.0], el)
        with M.Mlp(seed=2) as fresh:
            fresh.set_weights(w2_dev)
            out2_fresh = fresh.eval(cfg, [0.0, 0.0, 0.0], el)
    z = np.zeros_like(w0, dtype=np.float32)
    w1, m1, v1 = _adam_f32(w0, g0, z, z, 1, batch)
    w2, _, _ = _adam_f32(w1_dev, g1, m1, v1, 2, batch)
    # the first Adam step moves every parameter whose gradient is well above
    # eps by lr (m/sqrt(v) = sign(g) at step 1)
    moved = np.abs(g0) / batch > 1e-5
    assert moved.sum() > 0.5 * w0.size
    assert np.allclose(np.abs(w1_dev - w0)[moved], 1e-3, rtol=1e-2)
    assert np.max(np.abs(w1_dev - w1)) <= 1e-6 * 1e-3 + 4 * np.spacing(np.abs(w0).astype(np.float32)).max()
    assert np.max(np.abs(w2_dev - w2)) <= 1e-6 * 1e-3 + 4 * np.spacing(np.abs(w1_dev)).max()
    # a wrong bias correction (none at step 2) would miss by far more
    w2_wrong = w1_dev - np.float32(1e-3) * (np.float32(0.9) * m1 + np.float32(0.1) * (-g1 / batch)) / (
        np.sqrt(np.float32(0.999) * v1 + np.float32(0.001) * (g1 / batch) ** 2) + np.float32(1e-8))
    assert np.max(np.abs(w2_dev - w2_wrong)) > 100 * np.max(np.abs(w2_dev - w2))
    # the bf16 fragments were rebuilt from the updated master weights
    assert not np.array_equal(out2, out0)
    assert np.array_equal(out2, out2_fresh)


@pytest.fixture(scope="module")
def ne_learned_record():
    """The learned-proposal workload on the Ne atom (1s/2s/2p STO basis,
    seed 1): 1M box-sampler configurations, 8,000 Adam steps of the log-ESJD
    objective on the device, then the learned proposal and four box sides run
    from the same ensemble for 400 sampled sweeps each."""
    from mole_amd import workloads as WL

    return WL.learned(Z=10, walkers=32768, n_configs=1_000_000, sweeps=400, block=50,
                      box_sides=(0.1, 0.15, 0.2, 0.3), objective=M.MLP_OBJ_LOG_ESJD, log=lambda r: None)


def test_learned_energy_agrees_with_box_on_ne(ne_learned_record):
    # the exact MH ratio keeps psi^2 stationary: the learned chain's mean
    # energy agrees with the box chains' (errors from tau_int and the
    # variance; the box sides themselves must agree too)
    r = ne_learned_record
    sig = lambda x: max(x["error"], x["error_tau"])
    lrn = r["learned"]
    assert lrn["energy"] < 0.0 and r["learned"]["acceptance"] > 0.3
    for b in r["box"]:
        d = lrn["energy"] - b["energy"]
        assert abs(d) < 4.0 * np.hypot(sig(lrn), sig(b)), (b["step"], lrn["energy"], b["energy"], sig(lrn), sig(b))


def test_trained_network_lowers_tau_int_below_every_box_side_on_ne(ne_learned_record):
    # what the trained network gains over the method without it: tau_int of
    # E_L per sweep below the best box side's (tau_int by the reference's
    # statfor estimator on the device)
    r = ne_learned_record
    best_box_tau = min(b["tau_int_mean"] for b in r["box"])
    assert r["learned"]["tau_int_mean"] < 0.8 * best_box_tau, (r["learned"]["tau_int_mean"], best_box_tau)
    assert r["tau_ratio_vs_best_box"] > 1.25
    # the training objective rose
    assert r["training"]["objective_last"] > r["training"]["objective_first"] + 0.5


@pytest.mark.parametrize("Z", [3, 6, 8, 9, 10])
def test_features_match_oracle(Z):
    # the network inputs bit for bit against the oracle, for random
    # configurations and for ones whose partner distances tie exactly (the
    # lower index wins)
    system = S.atom(Z)
    wv = SJ.SjWave(system)
    N = Z
    rng = np.random.default_rng(Z)
    n = 300
    cfg = rng.normal(scale=0.9, size=(n, N, 3))
    el = rng.integers(0, N, n)
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=float)
    for t in range(24):
        # electron i at the origin, the rest on lattice points at distances
        # 1, 1, ..., 2, 2, ... in a shuffled order
        i = t % N
        others = [k for k in range(N) if k != i]
        rng.shuffle(others)
        c = np.zeros((N, 3))
        for m, k in enumerate(others):
            c[k] = dirs[m % 6] * (1 + m // 6 // (1 + t % 2))
        cfg[t] = c
        el[t] = i
    got = M.debug_mlp_features(cfg.reshape(n, -1), system.nuclei, el)
    want = np.array([ML.features(wv, cfg[e].ravel(), int(el[e])) for e in range(n)])
    assert np.array_equal(got[:, :13], want.astype(np.float32))
    assert np.all(got[:, 13:] == 0.0)
