"""Parity of the gfx950 two-electron walker path (through the C ABI) with the
CPU oracle on the same Philox streams: device primitives bit for bit, accept
decisions bit for bit, local energies within 1e-10, reductions within
rounding, and the reference's statistical anchors."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import mole_amd as M
from oracle import oracle_ffi as O

pytestmark = pytest.mark.gpu

E_TOL = 1e-10  # per-step local energy, FP64


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


# ---------------------------------------------------------------- primitives

def test_device_philox_matches_oracle():
    L = O.lib()
    for seed, sys_, w, ep, first in [(0, 0, 0, 0, 0), (0xDEADBEEF12345678, 3, 4095, 7, 12345),
                                      (42, 1, 77, 0, (1 << 33) + 1)]:
        got = M.debug_philox(seed, sys_, w, ep, first, 1000)
        want = np.array([L.orc_stream_u64(seed, sys_, w, ep, first + i) for i in range(1000)],
                        dtype=np.uint64)
        assert np.array_equal(got, want)


def test_device_exp_bitwise_matches_oracle():
    L = O.lib()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-40, 10, 50000), rng.uniform(-745, 709, 5000),
                        [0.0, -0.0, 1e-300, -1e-300, 709.7, -745.1, 800.0, -800.0]])
    got = M.debug_math(M.DEBUG_EXP, x)
    want = np.array([L.orc_exp(float(v)) for v in x])
    assert np.array_equal(bits(got), bits(want))


def test_device_sqrt_and_division_are_correctly_rounded():
    # numpy's sqrt and / are IEEE correctly rounded on the host
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.uniform(0, 10, 100000), 10.0 ** rng.uniform(-300, 300, 20000)])
    assert np.array_equal(bits(M.debug_math(M.DEBUG_SQRT, a)), bits(np.sqrt(a)))
    b = rng.uniform(-5, 5, a.size)
    assert np.array_equal(bits(M.debug_math(M.DEBUG_DIV, a, b)), bits(a / b))


def test_device_u64_maps_match_oracle():
    L = O.lib()
    rng = np.random.default_rng(2)
    u = rng.integers(0, 2**63, 20000, dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(1)
    u = np.concatenate([u, np.array([0, 2**64 - 1], dtype=np.uint64)])
    as_d = u.view(np.float64)
    got = M.debug_math(M.DEBUG_STANDARD, as_d)
    want = np.array([L.orc_u64_standard(int(v)) for v in u])
    assert np.array_equal(bits(got), bits(want))
    for half in (0.25, 1.0):
        got = M.debug_math(M.DEBUG_RANGE, as_d, np.full(u.size, half))
        want = np.array([L.orc_u64_range(int(v), -half, half) for v in u])
        assert np.array_equal(bits(got), bits(want))


# ---------------------------------------------------------------- trajectories

def _systems(kind, W, params):
    if kind == "he":
        return M.helium(W, *params), O.helium(*params)
    return M.h2(W, *params), O.h2(*params)


TRAJ_CASES = [
    # kind, params, walkers, steps, block
    ("he", (1.6875, 0.5, 0.3), 192, 80, 20),
    ("h2", (1.0, 0.5, 0.3), 192, 80, 20),
    ("he", (2.0, 0.0, 0.3), 192, 80, 20),
    ("h2", (1.2, 0.4, 0.6), 100, 57, 10),   # ragged: 100 walkers, 50 sweeps, sampling from sweep 10
    ("he", (1.6875, 0.5, 0.3), 1, 23, 7),   # one walker, 21 sweeps
]


@pytest.mark.parametrize("variant", [M.VARIANT_LANE, M.VARIANT_WAVESPEC, M.VARIANT_QUAD])
@pytest.mark.parametrize("kind,params,W,steps,B", TRAJ_CASES)
def test_trajectory_parity_with_oracle(kind, params, W, steps, B, variant):
    side, seed = 0.5, 0
    gsys, owf = _systems(kind, W, params)
    with M.Context([gsys], seed=seed, step=side, block_size=B, flags=M.FLAG_TRACE | M.FLAG_SAMPLES,
                   record_sweeps=steps) as ctx:
        ctx.set_launch_config(variant)
        ctx.init_walkers()
        ctx.run(steps)
        total = (steps // B) * B
        ns = total - B
        masks, el = ctx.trace(total, ns)
        cfg = ctx.get_configs()
        st = ctx.stats()
        e_s, psi_s, g_s = ctx.samples(ns)
    ref = O.run_box(owf, seed=seed, walkers=W, steps=steps, block_size=B, box_side=side, pgrad=True)
    assert np.array_equal(masks, ref["accept_mask"]), "accept/reject decisions differ"
    assert np.array_equal(bits(cfg), bits(ref["cfg"]))
    scale = np.maximum(1.0, np.abs(ref["e_local"]))
    assert np.max(np.abs(el - ref["e_local"]) / scale) <= E_TOL
    assert np.array_equal(e_s, el)
    assert np.max(np.abs(psi_s - ref["psi"]) / np.abs(ref["psi"])) <= 1e-13
    assert np.max(np.abs(g_s - ref["pgrad"]) / (1.0 + np.abs(ref["pgrad"]))) <= 1e-10
    # statistics
    assert st["n_accepted"] == ref["n_accepted"]
    assert st["n_moves"] == ref["n_moves"]
    assert st["n_samples"] == ref["n_samples"] and st["n_blocks"] == ref["n_blocks"]
    assert abs(st["sum_e"] - ref["sum_e"]) <= 1e-11 * ref["n_samples"]
    assert abs(st["sum_e2"] - ref["sum_e2"]) <= 1e-10 * ref["sum_e2"]
    assert abs(st["sum_block_mean"] - ref["sum_bm"]) <= 1e-11 * ref["n_blocks"]
    assert abs(st["sum_block_mean2"] - ref["sum_bm2"]) <= 1e-10 * ref["sum_bm2"]


def test_three_nuclei_general_path_parity():
    # H3+-like LCAO with three centres runs on the one-lane kernel only
    W, steps, B = 96, 40, 10
    nuc = [(0.0, 0.0, -1.0), (0.0, 0.9, 0.4), (0.0, -0.9, 0.4)]
    gsys = M.System(M.WF_LCAO2_PADE, nuc, [1, 1, 1], [1.1, 0.5, 0.3], W)
    owf = O.make_wf2(O.WF_LCAO2_PADE, (1.1, 0.5, 0.3), nuc, [1, 1, 1])
    with M.Context([gsys], seed=4, step=0.5, block_size=B, flags=M.FLAG_TRACE, record_sweeps=steps) as ctx:
        with pytest.raises(M.MoleError) as ei:
            ctx.set_launch_config(M.VARIANT_WAVESPEC)
        assert ei.value.status == M.ERR_UNSUPPORTED
        ctx.init_walkers()
        ctx.run(steps)
        masks, el = ctx.trace(steps, steps - B)
    ref = O.run_box(owf, seed=4, walkers=W, steps=steps, block_size=B, box_side=0.5)
    assert np.array_equal(masks, ref["accept_mask"])
    assert np.max(np.abs(el - ref["e_local"]) / np.maximum(1, np.abs(ref["e_local"]))) <= E_TOL


@pytest.mark.parametrize("make", [M.h2, M.helium], ids=["h2", "he"])
def test_variants_agree_on_split_sweeps(make):
    # arbitrary sweep splits (not multiples of the chunk), sampling switched
    # on mid-stream, a walker count that leaves a partial workgroup: every
    # kernel (one lane, wave-specialised, four lanes per walker) leaves
    # identical state and statistics
    res = []
    for variant in (M.VARIANT_LANE, M.VARIANT_WAVESPEC, M.VARIANT_QUAD):
        with M.Context([make(130)], seed=8, step=0.5, block_size=6) as ctx:
            ctx.set_launch_config(variant)
            ctx.init_walkers()
            ctx.sweep(7)
            ctx.sweep(13, sample=True)
            ctx.sweep(5, sample=True)
            res.append((ctx.get_configs(), ctx.stats(), ctx.sr_sums()))
    c0, s0, r0 = res[0]
    for c1, s1, r1 in res[1:]:
        assert np.array_equal(bits(c0), bits(c1))
        assert s0["n_accepted"] == s1["n_accepted"]
        for k in ("sum_e", "sum_e2", "sum_block_mean", "sum_block_mean2"):
            assert s0[k] == s1[k], k
        for k in ("sum_o", "sum_oe", "sum_oo"):
            assert np.array_equal(r0[k], r1[k]), k


@pytest.mark.parametrize("conv", [M.O_LOGDERIV, M.O_REFERENCE])
def test_sr_sums_match_oracle_samples(conv):
    W, steps, B, seed = 128, 60, 20, 5
    gsys, owf = _systems("h2", W, (1.1, 0.45, 0.35))
    with M.Context([gsys], seed=seed, step=0.5, block_size=B, o_convention=conv) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        sums = ctx.sr_sums()
    ref = O.run_box(owf, seed=seed, walkers=W, steps=steps, block_size=B, box_side=0.5, pgrad=True,
                    o_convention=conv)
    Ov = (ref["pgrad"] / ref["psi"][..., None]).reshape(-1, 3)
    E = ref["e_local"].ravel()
    assert sums["n"] == E.size
    def close(a, b, scale):
        assert np.all(np.abs(np.asarray(a) - np.asarray(b)) <= 1e-9 * scale), (a, b)
    close(sums["sum_e"], E.sum(), np.abs(E).sum())
    close(sums["sum_o"], Ov.sum(0), np.abs(Ov).sum(0))
    close(sums["sum_oe"], (Ov * E[:, None]).sum(0), np.abs(Ov * E[:, None]).sum(0))
    close(sums["sum_oo"], Ov.T @ Ov, np.abs(Ov).T @ np.abs(Ov))


def test_eval_matches_oracle():
    L = O.lib()
    rng = np.random.default_rng(3)
    X = rng.uniform(-2, 2, size=(500, 6))
    for kind, params in [("he", (1.6875, 0.5, 0.3)), ("h2", (1.0, 0.5, 0.3))]:
        gsys, owf = _systems(kind, 4, params)
        with M.Context([gsys]) as ctx:
            out = ctx.eval(X)
        for i, x in enumerate(X):
            xa = np.ascontiguousarray(x)
            v = L.orc_wf2_value(C.byref(owf), O.dptr(xa))
            assert bits(out["value"][i]) == bits(v)
            g = np.zeros(6)
            L.orc_wf2_gradient(C.byref(owf), O.dptr(xa), O.dptr(g))
            assert np.allclose(out["grad"][i], g, rtol=1e-11, atol=1e-13 * abs(v))
            lp = L.orc_wf2_laplacian(C.byref(owf), O.dptr(xa))
            assert abs(out["lap"][i] - lp) <= 1e-11 * max(abs(lp), abs(v))
            e = L.orc_wf2_local_energy(C.byref(owf), O.dptr(xa))
            assert abs(out["e_local"][i] - e) <= E_TOL * max(1.0, abs(e))
            pg = np.zeros(3)
            L.orc_wf2_parameter_gradient(C.byref(owf), O.dptr(xa), O.dptr(pg))
            assert np.allclose(out["pgrad"][i], pg, rtol=1e-11, atol=1e-14)


def test_continuation_checkpoint_bitwise():
    W = 64
    with M.Context([M.helium(W)], seed=3, step=0.5, block_size=10) as a:
        a.init_walkers()
        a.sweep(50)
        full = a.get_configs()
    with M.Context([M.helium(W)], seed=3, step=0.5, block_size=10) as b:
        b.init_walkers()
        b.sweep(20)
        cfg, (ep, pos) = b.get_configs(), b.get_stream()
    with M.Context([M.helium(W)], seed=3, step=0.5, block_size=10) as c:
        c.set_configs(cfg)
        c.set_stream(ep, pos)
        c.sweep(30)
        assert np.array_equal(bits(c.get_configs()), bits(full))


def test_run_asserts_two_blocks():
    with M.Context([M.helium(8)], block_size=10) as ctx:
        with pytest.raises(M.MoleError) as ei:
            ctx.run(19)
        assert ei.value.status == M.ERR_INVALID_ARGUMENT
        with pytest.raises(M.MoleError):
            ctx.run(40, 20)


def test_vmc_reference_schedule_parity():
    """VmcRunner::run_optimization (vmc.rs:43-106) in its literal schedule:
    every iteration restarts all workers from the master configuration with
    fresh streams; the device run must follow the oracle's trajectory."""
    iters, W, steps, B, seed = 3, 96, 60, 20, 11
    params = [1.5, 0.5, 0.3]
    owf = O.h2(*params)
    op = O.OptParams(optimizer=O.OPT_SR, step=0.05, sr_mode=O.SR_TEXTBOOK, sr_eps=1e-3,
                     o_convention=O.O_LOGDERIV, persistent=0)
    en, er, ac, gr, pa = (np.zeros(iters), np.zeros(iters), np.zeros(iters), np.zeros(iters * 3),
                          np.zeros(iters * 3))
    rc = O.lib().orc_vmc_optimize(C.byref(owf), seed, 0.5, iters, W * steps, B, W, C.byref(op), 4,
                                  O.dptr(en), O.dptr(er), O.dptr(ac), O.dptr(gr), O.dptr(pa))
    assert rc == 0
    master = np.array([O.lib().orc_u64_range(O.lib().orc_stream_u64(seed, 0, 0xFFFFFFFF, 0, k), -1.0, 1.0)
                       for k in range(6)])
    with M.Context([M.h2(W, *params)], seed=seed, step=0.5, block_size=B) as ctx:
        for it in range(iters):
            ctx.broadcast_config(master)
            ctx.set_stream(it + 1, 0)
            ctx.reset_stats()
            ctx.run(steps)
            st = ctx.stats()
            raw = ctx.sr_sums_raw()
            dp, g = M.parameter_update(raw, M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
            mean = st["sum_e"] / st["n_samples"]
            assert abs(mean - en[it]) <= 1e-12 * abs(en[it])
            nb = st["n_blocks"]
            err = np.sqrt((st["sum_block_mean2"] / nb - mean**2) / (nb - 1))
            assert abs(err - er[it]) <= 1e-6 * er[it]
            assert abs(st["n_accepted"] / 2 / (W * steps) - ac[it]) <= 1e-15
            assert np.allclose(g, gr[3 * it:3 * it + 3], rtol=1e-9, atol=1e-12)
            p = ctx.get_params() + dp
            ctx.set_params(p)
            assert np.allclose(p, pa[3 * it:3 * it + 3], rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- physics anchors

@pytest.mark.parametrize("zeta", [1.5, 27.0 / 16.0])
def test_helium_product_sto_energy_anchor(zeta):
    # tests/helium_lcao.rs:134: <E>(zeta) = zeta^2 - 27 zeta/8 for the product STO
    with M.Context([M.helium(4096, zeta, 0.0, 0.3)], seed=21, step=1.0, block_size=100) as ctx:
        ctx.init_walkers()
        ctx.run(2000)
        st = ctx.stats()
    mean = st["sum_e"] / st["n_samples"]
    nb = st["n_blocks"]
    err = np.sqrt((st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2) / (nb - 1))
    exact = zeta**2 - 27 * zeta / 8
    assert abs(mean - exact) < 4 * err, (mean, exact, err)


def test_helium_zero_variance_ionic():
    # zeta = Z = 2, no Jastrow, IonicHamiltonian: E_L == -4 on every sample
    with M.Context([M.helium(1024, 2.0, 0.0, 0.0, hamiltonian=M.H_IONIC)], seed=2, step=0.5,
                   block_size=10) as ctx:
        ctx.init_walkers()
        ctx.run(100)
        st = ctx.stats()
    mean = st["sum_e"] / st["n_samples"]
    assert abs(mean + 4.0) < 1e-12
    assert st["sum_e2"] / st["n_samples"] - mean**2 < 1e-10


def test_uniform_psi_always_accepts():
    with M.Context([M.helium(256, 0.0, 0.0, 0.0)], seed=1, step=1.0, block_size=10) as ctx:
        ctx.init_walkers()
        ctx.run(40)
        st = ctx.stats()
    assert st["n_accepted"] == st["n_moves"] == 256 * 40 * 2


def test_device_optimisation_loop_equals_host_loop():
    # mole_optimize keeps the SR loop on the device; the trajectory must be
    # the host loop's bit for bit (same sums, same solve, same rounding)
    W, sweeps, block, iters = 1024, 60, 10, 5
    p0 = np.array([1.0, 0.5, 0.3])
    host_p, host_e = p0.copy(), []
    with M.Context([M.h2(W, *p0)], seed=7, step=0.5, block_size=block) as ctx:
        ctx.init_walkers()
        for _ in range(iters):
            ctx.reset_stats()
            ctx.run(sweeps, block)
            st = ctx.stats()
            host_e.append((st["sum_e"], st["n_samples"], st["n_accepted"]))
            dp, _ = M.parameter_update(ctx.sr_sums_raw(), M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
            host_p = host_p + dp
            ctx.set_params(host_p)
        cfg_host = ctx.get_configs()
    with M.Context([M.h2(W, *p0)], seed=7, step=0.5, block_size=block) as ctx:
        ctx.init_walkers()
        p, log = ctx.optimize(iters, sweeps, block, M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
        assert np.array_equal(ctx.get_params(), p)
        cfg_dev = ctx.get_configs()
        # the context continues from the device loop's parameters and walkers
        ctx.reset_stats()
        ctx.run(sweeps, block)
    assert np.array_equal(p, host_p)
    for k, (se, n, acc) in enumerate(host_e):
        assert log[k, 0] == se and log[k, 6] == n and log[k, 4] == acc and log[k, 5] == 0.0
    assert np.array_equal(cfg_dev.view(np.uint64), cfg_host.view(np.uint64))


def test_helium_energy_run_full_shape():
    # He (zeta 1.6875, Pade b 0.5, c 0.3), 4,096 walkers x 10,000 sweeps,
    # block 100 (the first discarded), box side 0.5: one Runner::run whose
    # final configurations are the oracle's bit for bit (so are all 8.2e7
    # accept decisions on the way), with the same sums
    W, S, B, side = 4096, 10000, 100, 0.5
    print(f"helium energy run W={W} sweeps={S}", flush=True)
    with M.Context([M.helium(W, 1.6875, 0.5, 0.3)], seed=0, step=side, block_size=B) as ctx:
        ctx.init_walkers()
        ctx.run(S)
        cfg = ctx.get_configs()
        st = ctx.stats()
    ref = O.run_box(O.helium(1.6875, 0.5, 0.3), seed=0, walkers=W, steps=S, block_size=B, box_side=side,
                    record=False, n_threads=8)
    assert np.array_equal(bits(cfg), bits(ref["cfg"]))
    assert st["n_accepted"] == ref["n_accepted"] and st["n_moves"] == ref["n_moves"]
    assert st["n_samples"] == ref["n_samples"] and st["n_blocks"] == ref["n_blocks"]
    assert abs(st["sum_e"] - ref["sum_e"]) <= 1e-12 * abs(ref["sum_e"])
    assert abs(st["sum_block_mean2"] - ref["sum_bm2"]) <= 1e-12 * ref["sum_bm2"]


def test_benchmark_shape_parity_and_device_loop():
    # the H2 + SR benchmark at its full shape: 16,384 walkers x 200 sweeps,
    # block 20, box side 0.5, parameters (1.0, 0.5, 0.3).  One Runner::run
    # against the oracle (every accept decision, the final configurations,
    # the sums), then three SR iterations (step 0.05, shift 1e-3) of the
    # device loop against the host loop, bit for bit.
    W, S, B, side, p0 = 16384, 200, 20, 0.5, (1.0, 0.5, 0.3)
    print(f"benchmark shape W={W} sweeps={S}", flush=True)
    with M.Context([M.h2(W, *p0)], seed=0, step=side, block_size=B, flags=M.FLAG_TRACE,
                   record_sweeps=S) as ctx:
        ctx.init_walkers()
        ctx.run(S)
        masks, el = ctx.trace(S, S - B)
        cfg = ctx.get_configs()
        st = ctx.stats()
    ref = O.run_box(O.h2(*p0), seed=0, walkers=W, steps=S, block_size=B, box_side=side)
    assert np.array_equal(masks, ref["accept_mask"]), "accept/reject decisions differ"
    assert np.array_equal(bits(cfg), bits(ref["cfg"]))
    assert st["n_accepted"] == ref["n_accepted"] and st["n_samples"] == ref["n_samples"]
    scale = np.maximum(1.0, np.abs(ref["e_local"]))
    assert np.max(np.abs(el - ref["e_local"]) / scale) <= E_TOL
    assert abs(st["sum_e"] - ref["sum_e"]) <= 1e-11 * ref["n_samples"]

    iters = 3
    host_p = np.array(p0)
    with M.Context([M.h2(W, *p0)], seed=7, step=side, block_size=B) as ctx:
        ctx.init_walkers()
        for _ in range(iters):
            ctx.reset_stats()
            ctx.run(S, B)
            dp, _ = M.parameter_update(ctx.sr_sums_raw(), M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
            host_p = host_p + dp
            ctx.set_params(host_p)
        cfg_host = ctx.get_configs()
    with M.Context([M.h2(W, *p0)], seed=7, step=side, block_size=B) as ctx:
        ctx.init_walkers()
        p, log = ctx.optimize(iters, S, B, M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
        cfg_dev = ctx.get_configs()
    assert np.all(log[:, 5] == 0.0)
    assert np.array_equal(p.view(np.uint64), host_p.view(np.uint64))
    assert np.array_equal(cfg_dev.view(np.uint64), cfg_host.view(np.uint64))


def test_benchmark_sr_trajectory_matches_oracle():
    # the benchmark's whole optimisation: 50 SR iterations of 16,384
    # persistent walkers x 200 sweeps (block 20, box side 0.5, step 0.05,
    # shift 1e-3, parameters from (1.0, 0.5, 0.3), seed 0) on the device
    # loop, against the oracle's VmcRunner with persistent walkers (two full
    # psi evaluations per move, gradient and SR matrix from the stored
    # samples) on the host cores.  Accept decisions are bit-exact; the sums
    # are formed in different orders, so energies and parameters agree to
    # rounding, iteration by iteration.
    iters, W, S, B, side, p0, seed = 50, 16384, 200, 20, 0.5, (1.0, 0.5, 0.3), 0
    print(f"benchmark SR trajectory: {iters} iterations of W={W} x {S} sweeps", flush=True)
    owf = O.h2(*p0)
    op = O.OptParams(optimizer=O.OPT_SR, step=0.05, sr_mode=O.SR_TEXTBOOK, sr_eps=1e-3,
                     o_convention=O.O_LOGDERIV, persistent=1)
    en, er, ac, gr, pa = (np.zeros(iters), np.zeros(iters), np.zeros(iters), np.zeros(iters * 3),
                          np.zeros(iters * 3))
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rc = O.lib().orc_vmc_optimize(C.byref(owf), seed, side, iters, W * S, B, W, C.byref(op), threads,
                                  O.dptr(en), O.dptr(er), O.dptr(ac), O.dptr(gr), O.dptr(pa))
    assert rc == 0
    with M.Context([M.h2(W, *p0)], seed=seed, step=side, block_size=B) as ctx:
        ctx.init_walkers()
        p, log = ctx.optimize(iters, S, B, M.OPT_SR, 0.05, M.SR_TEXTBOOK, 1e-3)
    assert np.all(log[:, 5] == 0.0)
    e_dev = log[:, 0] / log[:, 6]
    assert np.allclose(e_dev, en, rtol=1e-11, atol=0.0), np.max(np.abs(e_dev - en))
    assert np.allclose(log[:, 4] / (2.0 * W * S), ac, rtol=0.0, atol=1e-15)
    assert np.allclose(p, pa[-3:], rtol=1e-9, atol=0.0), (p, pa[-3:])
    # the parameters moved (zeta from 1.0 to ~1.26)
    assert abs(pa[-3] - 1.0) > 0.1


@pytest.mark.parametrize("kind,kw", [(M.OPT_MOMENTUM, dict(momentum=0.5)), (M.OPT_NESTEROV, dict(momentum=0.5)),
                                     (M.OPT_LBFGS, dict(history=3))])
def test_device_loop_stateful_optimizers_equal_host_loop(kind, kw):
    # MomentumDescent / NesterovMomentum / OnlineLbfgs (optimizers.rs:32-179)
    # advanced on the device by mole_optimize_with: parameters, walkers and
    # the optimizer state after the loop equal the host loop's bit for bit
    W, sweeps, block, iters, step = 512, 40, 10, 4, 0.01
    p0 = np.array([1.0, 0.5, 0.3])
    host_opt = M.Optimizer(kind, 3, step, **kw)
    host_p = p0.copy()
    with M.Context([M.h2(W, *p0)], seed=5, step=0.5, block_size=block) as ctx:
        ctx.init_walkers()
        for _ in range(iters):
            ctx.reset_stats()
            ctx.run(sweeps, block)
            dp, _ = host_opt.compute_parameter_update(host_p, ctx.sr_sums_raw())
            host_p = host_p + dp
            ctx.set_params(host_p)
        cfg_host = ctx.get_configs()
    dev_opt = M.Optimizer(kind, 3, step, **kw)
    with M.Context([M.h2(W, *p0)], seed=5, step=0.5, block_size=block) as ctx:
        ctx.init_walkers()
        p, log = ctx.optimize_with(dev_opt, iters, sweeps, block)
        cfg_dev = ctx.get_configs()
    assert np.all(np.isfinite(p)) and np.all(log[:, 5] == 0.0)
    assert np.array_equal(p.view(np.uint64), host_p.view(np.uint64))
    assert np.array_equal(cfg_dev.view(np.uint64), cfg_host.view(np.uint64))
    assert bytes(dev_opt.state) == bytes(host_opt.state)
