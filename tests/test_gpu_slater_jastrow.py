"""Parity of the gfx950 Slater-Jastrow path (through the C ABI) with the CPU
oracle's Sherman-Morrison runner on the same Philox streams, in FP64 and in
mixed precision; independence of the result from lanes-per-walker;
heterogeneous systems in one launch; zero-variance anchors (1s and 2p); the
FP64 log-determinants carried through the updates and their drift; the
singular-pivot guard on planted singular configurations."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import systems as S
from oracle import sj_ffi as SJ

pytestmark = pytest.mark.gpu

E_TOL = 1e-10


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def run_device(systems, W, steps, B, *, precision=M.PREC_FP64, lu=16, lanes=None, seed=0, side=0.6,
               hamiltonian=M.H_ELECTRONIC, proposal=M.PROPOSAL_BOX):
    dev = [M.slater_jastrow(s, W, hamiltonian=hamiltonian) for s in systems]
    with M.Context(dev, seed=seed, step=side, block_size=B, flags=M.FLAG_TRACE, record_sweeps=steps,
                   precision=precision, lu_period=lu, proposal=proposal) as ctx:
        if lanes is not None:
            for k in range(len(systems)):
                ctx.set_walker_lanes(lanes, sys=k)
        ctx.init_walkers()
        ctx.run(steps)
        total = (steps // B) * B
        out = []
        for k in range(len(systems)):
            masks, el = ctx.trace(total, total - B, sys=k)
            out.append(dict(masks=masks, el=el, cfg=ctx.get_configs(sys=k), stats=ctx.stats(sys=k),
                            sr=ctx.sr_sums(sys=k), lanes=ctx.walker_lanes(sys=k)))
    return out


def run_oracle(system, W, steps, B, *, precision=SJ.PREC_FP64, lu=16, seed=0, side=0.6, sys=0,
               hamiltonian=0, proposal=SJ.BOX):
    wv = SJ.SjWave(system, hamiltonian=hamiltonian)
    return SJ.run(wv, seed=seed, walkers=W, steps=steps, block_size=B, box_side=side,
                  precision=precision, lu_period=lu, sys=sys, n_threads=8, proposal=proposal, tau=side)


@pytest.mark.parametrize("Z", [3, 6, 10])
def test_fp64_parity_with_oracle(Z):
    W, steps, B = 70, 30, 10
    system = S.atom(Z)
    d = run_device([system], W, steps, B)[0]
    o = run_oracle(system, W, steps, B)
    assert np.array_equal(d["masks"], o["accept_mask"]), "accept/reject decisions differ"
    assert np.array_equal(bits(d["cfg"]), bits(o["cfg"]))
    rel = np.abs(d["el"] - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= E_TOL
    st = d["stats"]
    assert st["n_accepted"] == o["n_accepted"] and st["n_moves"] == o["n_moves"]
    assert abs(st["sum_e"] - o["sum_e"]) <= 1e-10 * o["n_samples"] * max(1.0, abs(o["sum_e"] / o["n_samples"]))
    assert st["max_condition"] > 1.0


@pytest.mark.parametrize("Z", [3, 6])
def test_mixed_precision_parity_with_oracle(Z):
    W, steps, B = 64, 30, 10
    system = S.atom(Z)
    d = run_device([system], W, steps, B, precision=M.PREC_MIXED, lu=64)[0]
    o = run_oracle(system, W, steps, B, precision=SJ.PREC_MIXED, lu=64)
    assert np.array_equal(d["masks"], o["accept_mask"])
    assert np.array_equal(bits(d["cfg"]), bits(o["cfg"]))
    rel = np.abs(d["el"] - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= E_TOL


@pytest.mark.parametrize("Z,lanes,steps", [(3, 4, 30), (4, 4, 12), (5, 4, 10), (6, 4, 10), (7, 4, 8), (8, 8, 8),
                                           (9, 8, 8), (10, 8, 10)])
def test_mixed_precision_parity_at_workload_width(Z, lanes, steps):
    # the atoms workload's width (65,536 walkers, LU refresh every 64 moves)
    # with the lanes per walker its autotuning picks: every accept decision,
    # the configurations and the FP32 inverses' consequences bit for bit
    W, B = 65536, 2
    print(f"workload width Z={Z} W={W} lanes={lanes} sweeps={steps}", flush=True)
    system = S.atom(Z)
    d = run_device([system], W, steps, B, precision=M.PREC_MIXED, lu=64, lanes=lanes, seed=1)[0]
    o = run_oracle(system, W, steps, B, precision=SJ.PREC_MIXED, lu=64, seed=1)
    assert d["lanes"] == lanes
    assert np.array_equal(d["masks"], o["accept_mask"]), "accept/reject decisions differ"
    assert np.array_equal(bits(d["cfg"]), bits(o["cfg"]))
    rel = np.abs(d["el"] - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= 1e-9
    assert d["stats"]["n_accepted"] == o["n_accepted"]


def tile_footprint(system, lanes):
    """LDS bytes of one full tile of the system at the given lanes per walker
    (a one-tile context's launch shape)."""
    with M.Context([M.slater_jastrow(system, 64 // lanes)], seed=0, step=0.6, block_size=2,
                   precision=M.PREC_MIXED, lu_period=64) as c:
        c.set_walker_lanes(lanes)
        return c.launch_info()[1]


def test_heterogeneous_batch_at_workload_shape():
    # the heterogeneous workload's launches: 256 systems of 2..40 electrons,
    # 2^20 walkers split in proportion to N, the lanes its autotuning picked,
    # mixed precision.  The tiles are grouped into LDS footprint classes (4 KiB
    # steps), one launch per class with blockIdx.x restarting at 0 in each,
    # longest tiles first within a class.  The system with the fewest
    # electrons, the one with the largest tile footprint (system 52, 11
    # electrons) and the largest system of a third footprint class against
    # their own oracle runs (systems are independent streams)
    from mole_amd import autotune
    from mole_amd import workloads as WL
    lanes = {2: 4, 3: 4, 4: 4, 5: 8, 6: 8, 7: 8, 8: 8, 9: 16, 10: 8, 11: 8, 12: 16, 13: 16, 14: 16, 15: 16,
             16: 16, 17: 16, 18: 16, 19: 32, 20: 16}
    lanes.update({n: 32 for n in range(21, 32)})
    lanes.update({n: 64 for n in range(32, 41)})
    sys_list = WL.hetero_systems(256, 3)
    Ns = np.array([s.n_electrons for s in sys_list])
    W = np.maximum(1, np.floor(Ns / Ns.sum() * (1 << 20))).astype(int)
    W[np.argmax(Ns)] += (1 << 20) - W.sum()
    order = np.argsort(Ns, kind="stable")
    cls = lambda k: (tile_footprint(sys_list[k], lanes[int(Ns[k])]) + 4095) // 4096
    picks = [int(order[0]), 52]
    # the third: the most electrons among the systems of a third footprint
    # class, so class boundaries stay covered
    taken = {cls(k) for k in picks}
    for k in order[::-1]:
        if cls(int(k)) not in taken:
            picks.append(int(k))
            break
    assert len(picks) == 3 and len(taken | {cls(picks[2])}) == 3
    steps, B = 4, 2
    print(f"heterogeneous batch: {len(sys_list)} systems, {int(W.sum())} walkers, checking {picks}", flush=True)
    dev = [M.slater_jastrow(s, int(w)) for s, w in zip(sys_list, W)]
    with M.Context(dev, seed=3, step=0.6, block_size=B, flags=M.FLAG_TRACE, record_sweeps=steps,
                   precision=M.PREC_MIXED, lu_period=64) as ctx:
        autotune.apply(ctx, sys_list, lanes)
        ctx.init_walkers()
        ctx.run(steps)
        got = {k: (ctx.trace(steps, steps - B, sys=k), ctx.get_configs(sys=k)) for k in picks}
    for k in picks:
        (masks, el), cfg = got[k]
        o = SJ.run(SJ.SjWave(sys_list[k]), seed=3, walkers=int(W[k]), steps=steps, block_size=B, box_side=0.6,
                   precision=SJ.PREC_MIXED, lu_period=64, sys=k, n_threads=8)
        assert np.array_equal(masks, o["accept_mask"]), f"system {k}: accept/reject decisions differ"
        assert np.array_equal(bits(cfg), bits(o["cfg"])), f"system {k}"
        rel = np.abs(el - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
        assert np.max(rel) <= 1e-9, f"system {k}"


def test_results_independent_of_lanes_per_walker():
    system = S.atom(7)
    ref = None
    for L in (2, 4, 8, 16, 32, 64):
        d = run_device([system], 40, 20, 5, lanes=L)[0]
        assert d["lanes"] == L
        if ref is None:
            ref = d
            continue
        assert np.array_equal(d["masks"], ref["masks"]), L
        assert np.array_equal(bits(d["cfg"]), bits(ref["cfg"])), L
        assert np.array_equal(bits(d["el"]), bits(ref["el"])), L
        for k in ("sum_o", "sum_oe", "sum_oo"):
            assert np.array_equal(d["sr"][k], ref["sr"][k]), (L, k)


def test_heterogeneous_systems_in_one_launch():
    systems = [S.helium_product(1.7, 0.5), S.atom(4), S.hydrogen_chain(4), S.random_molecule(9, seed=2),
               S.atom(10)]
    W, steps, B = 33, 20, 5
    dev = run_device(systems, W, steps, B)
    for k, system in enumerate(systems):
        o = run_oracle(system, W, steps, B, sys=k)
        assert np.array_equal(dev[k]["masks"], o["accept_mask"]), k
        assert np.array_equal(bits(dev[k]["cfg"]), bits(o["cfg"])), k
        rel = np.abs(dev[k]["el"] - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
        assert np.max(rel) <= E_TOL, k


def test_zero_variance_anchors():
    d = run_device([S.hydrogenic(1)], 128, 20, 5)[0]
    assert np.max(np.abs(d["el"] + 0.5)) <= 1e-12
    # the 2p eigenfunctions x_q exp(-Z r/2): E_L = -Z^2/8 for every sample
    for axis in range(3):
        d = run_device([S.hydrogenic_2p(2, axis)], 128, 20, 5)[0]
        assert np.max(np.abs(d["el"] + 0.5)) <= 1e-11
    d = run_device([S.helium_product(2.0, 0.0)], 128, 20, 5, hamiltonian=M.H_IONIC)[0]
    assert np.max(np.abs(d["el"] + 4.0)) <= 1e-12


def test_helium_product_anchor_through_determinants():
    zeta = 1.5
    sysd = M.slater_jastrow(S.helium_product(zeta, 0.0), 4096)
    with M.Context([sysd], seed=9, step=1.0, block_size=50, lu_period=32) as ctx:
        ctx.init_walkers()
        ctx.run(1000)
        st = ctx.stats()
    mean = st["sum_e"] / st["n_samples"]
    nb = st["n_blocks"]
    err = np.sqrt((st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (zeta**2 - 27 * zeta / 8)) < 4 * err


def test_graph_replays_match_eager_sweeps():
    # a captured [sweep, advance] launch replayed several times, with eager
    # work and synchronisations between replays, equals the eager sequence
    systems = [S.atom(4), S.random_molecule(6, seed=4)]

    def make():
        dev = [M.slater_jastrow(s, 50) for s in systems]
        ctx = M.Context(dev, seed=5, step=0.6, block_size=4, precision=M.PREC_MIXED, lu_period=16)
        ctx.init_walkers()
        return ctx

    snaps = []
    for use_graph in (False, True):
        ctx = make()
        g = ctx.capture(5, sample=True) if use_graph else None
        out = []
        for r in range(4):
            if use_graph:
                g.launch(1)
            else:
                ctx.sweep(5, sample=True)
            ctx.synchronize()
            ctx.sweep(2)  # eager work between replays
            out.append((ctx.get_configs(0), ctx.get_configs(1), ctx.stats(0), ctx.sr_sums(1), ctx.get_stream()))
        if g is not None:
            g.launch(2)
            out.append((ctx.get_configs(0), ctx.get_configs(1), ctx.stats(0), ctx.sr_sums(1), ctx.get_stream()))
            g.close()
        else:
            ctx.sweep(5, sample=True)
            ctx.sweep(5, sample=True)
            out.append((ctx.get_configs(0), ctx.get_configs(1), ctx.stats(0), ctx.sr_sums(1), ctx.get_stream()))
        ctx.close()
        snaps.append(out)
    for a, b in zip(*snaps):
        assert np.array_equal(bits(a[0]), bits(b[0]))
        assert np.array_equal(bits(a[1]), bits(b[1]))
        assert a[2] == b[2]
        for k in ("sum_o", "sum_oe", "sum_oo"):
            assert np.array_equal(a[3][k], b[3][k])
        assert a[4] == b[4]


def test_graph_goes_stale_when_lanes_change():
    dev = [M.slater_jastrow(S.atom(3), 32)]
    with M.Context(dev, seed=1, step=0.6, block_size=4, lu_period=16) as ctx:
        ctx.init_walkers()
        g = ctx.capture(2)
        g.launch(1)
        ctx.set_walker_lanes(2)
        with pytest.raises(M.MoleError) as ei:
            g.launch(1)
        assert ei.value.status == M.ERR_INVALID_ARGUMENT
        g.close()


# ------------------------------------------------ drift-diffusion proposal

@pytest.mark.parametrize("Z,precision", [(3, M.PREC_FP64), (7, M.PREC_FP64), (4, M.PREC_MIXED)])
def test_diffusion_parity_with_oracle(Z, precision):
    # MetropolisDiffuse (metrop.rs:150-197, single-row Green's functions):
    # same Box-Muller normals, drift and accept decisions as the oracle
    W, steps, B, tau = 48, 24, 8, 0.05
    system = S.atom(Z)
    d = run_device([system], W, steps, B, precision=precision, lu=16, side=tau,
                   proposal=M.PROPOSAL_DIFFUSE)[0]
    o = run_oracle(system, W, steps, B, precision=precision, lu=16, side=tau, proposal=SJ.DIFFUSE)
    assert np.array_equal(d["masks"], o["accept_mask"]), "accept/reject decisions differ"
    assert np.array_equal(bits(d["cfg"]), bits(o["cfg"]))
    rel = np.abs(d["el"] - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= E_TOL


def test_diffusion_heterogeneous_and_lane_independent():
    systems = [S.helium_product(1.7, 0.5), S.random_molecule(6, seed=8)]
    ref = None
    for L in (2, 8, 32):
        dev = run_device(systems, 24, 12, 4, lanes=L, side=0.1, proposal=M.PROPOSAL_DIFFUSE)
        if ref is None:
            ref = dev
            for k, system in enumerate(systems):
                o = run_oracle(system, 24, 12, 4, side=0.1, sys=k, proposal=SJ.DIFFUSE)
                assert np.array_equal(dev[k]["masks"], o["accept_mask"]), k
                assert np.array_equal(bits(dev[k]["cfg"]), bits(o["cfg"])), k
        else:
            for k in range(len(systems)):
                assert np.array_equal(bits(dev[k]["cfg"]), bits(ref[k]["cfg"])), (L, k)
                assert np.array_equal(bits(dev[k]["el"]), bits(ref[k]["el"])), (L, k)


def test_diffusion_helium_product_anchor():
    zeta = 1.5
    sysd = M.slater_jastrow(S.helium_product(zeta, 0.0), 4096)
    with M.Context([sysd], seed=3, step=0.3, block_size=50, lu_period=32,
                   proposal=M.PROPOSAL_DIFFUSE) as ctx:
        ctx.init_walkers()
        ctx.run(1000)
        st = ctx.stats()
    mean = st["sum_e"] / st["n_samples"]
    nb = st["n_blocks"]
    err = np.sqrt((st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (zeta**2 - 27 * zeta / 8)) < 4 * err


# ------------------------------------------------ log-determinants, singular blocks

@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_logdet_carried_through_updates_matches_oracle(precision):
    # det A_s carried through every accepted move as mantissa x 2^exponent:
    # the drift against each re-inversion (the FP32 chain's error in mixed
    # precision) bit for bit as the oracle's, the final log-determinants as
    # the oracle carries them, and (FP64) as the LU of the final configuration
    W, steps, B = 64, 60, 20
    system = S.atom(9)
    dev = [M.slater_jastrow(system, W)]
    with M.Context(dev, seed=2, step=0.6, block_size=B, precision=precision, lu_period=64) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        st = ctx.stats()
        ld, sg = ctx.logdet()
        cfg = ctx.get_configs()
    wv = SJ.SjWave(system)
    o = SJ.run(wv, seed=2, walkers=W, steps=steps, block_size=B, box_side=0.6, precision=precision, lu_period=64,
               n_threads=8, record=False)
    assert np.array_equal(bits(cfg), bits(o["cfg"]))
    assert st["max_logdet_drift"] == o["max_ld_drift"]
    assert st["n_singular"] == 0 == o["n_singular"]
    want = np.array([[SJ.logdet_value(o["ld"][w, 2 * q], o["ld"][w, 2 * q + 1]) for q in (0, 1)] for w in range(W)])
    assert np.max(np.abs(ld - want)) <= 1e-12 * np.max(np.abs(want))
    assert np.array_equal(sg, np.sign(o["ld"][:, [0, 2]]))
    if precision == M.PREC_FP64:
        assert 0.0 < st["max_logdet_drift"] < 1e-10
        for w in range(W):
            for q in (0, 1):
                lu, s_lu = SJ.logdet(wv, cfg[w], q)
                assert abs(ld[w, q] - lu) < 1e-10 and sg[w, q] == s_lu
    else:
        assert st["max_logdet_drift"] > 1e-8


def test_planted_singular_configuration_is_counted():
    # walkers 0..7 start with spin-up electrons 0 and 1 at the same point
    # (identical rows, det A_up = 0): the re-inversion counts the block and
    # skips its division instead of writing Inf/NaN into the walker; the
    # guarded path then follows the oracle's bit for bit, and the other
    # walkers' determinants stay finite
    W, steps, B = 16, 20, 10
    system = S.atom(6)
    cfg = np.random.default_rng(7).normal(scale=1.2, size=(W, 3 * system.n_electrons))
    cfg[:8, 3:6] = cfg[:8, 0:3]
    dev = [M.slater_jastrow(system, W)]
    with M.Context(dev, seed=4, step=0.6, block_size=B, flags=M.FLAG_TRACE, record_sweeps=steps, lu_period=8) as ctx:
        ctx.set_configs(cfg)
        ld0, sg0 = ctx.logdet()
        n0 = ctx.stats()["n_singular"]
        ctx.run(steps)
        masks, _ = ctx.trace(steps, 0)
        st = ctx.stats()
        out = ctx.get_configs()
    assert n0 == 8
    assert np.all(ld0[:8, 0] == -np.inf) and np.all(sg0[:8, 0] == 0.0)
    assert np.all(np.isfinite(ld0[8:])) and np.all(np.isfinite(ld0[:, 1]))
    o = SJ.run(SJ.SjWave(system), seed=4, walkers=W, steps=steps, block_size=B, box_side=0.6, lu_period=8,
               cfg=cfg, n_threads=8)
    assert np.array_equal(masks, o["accept_mask"])
    assert np.array_equal(bits(out), bits(o["cfg"]))
    assert st["n_singular"] == o["n_singular"] >= 8
    assert np.all(np.isfinite(out))


# ------------------------------------------------ reference anchors

H2ION_E = -0.43720930232581723   # quadrature of <E> of the product function (R = 2.5)
H2ION_STD = 0.5052200457935744


def test_h2_ion_product_reference_anchor():
    # tests/hydrogen_molecular_ion_lcao.rs:100-142 on the device: H2+ at
    # R = 2.5 with psi = exp(-r_A) exp(-r_B) (a constant orbital times the e-n
    # Pade factor, a = 1, d = 0), MetropolisBox(1.0), Runner::run(.., 100);
    # the reference asserts |E + 0.565| < std of the samples.  The device's
    # short run also follows the oracle bit for bit, and its long run's mean
    # sits on the quadrature value within the blocking error.
    system = S.h2_ion_product()
    d = run_device([system], 64, 40, 10, side=1.0, seed=4)[0]
    o = run_oracle(system, 64, 40, 10, side=1.0, seed=4)
    assert np.array_equal(d["masks"], o["accept_mask"])
    assert np.array_equal(bits(d["cfg"]), bits(o["cfg"]))
    with M.Context([M.slater_jastrow(system, 8192)], seed=0, step=1.0, block_size=100, lu_period=64) as ctx:
        ctx.init_walkers()
        ctx.run(3000)
        st = ctx.stats()
    n = st["n_samples"]
    mean = st["sum_e"] / n
    std = np.sqrt(st["sum_e2"] / n - mean * mean)
    nb = st["n_blocks"]
    err = np.sqrt((st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2) / (nb - 1))
    assert abs(mean - (-0.565)) < std
    assert abs(mean - H2ION_E) < 4 * err, (mean, err)
    assert abs(std - H2ION_STD) < 0.02 * H2ION_STD


@pytest.mark.parametrize("workgroups", [0, 3])
def test_persistent_schedule_matches_class_launches(workgroups):
    # every tile of the context in one persistent launch (the tiles drawn from
    # a device ticket, longest first; 3 workgroups: each takes many tiles)
    # against one launch per footprint class: the same walkers, bit for bit,
    # eagerly and replayed from a captured graph (the ticket's reset is part
    # of the capture)
    systems = [S.atom(4), S.random_molecule(9, seed=2), S.hydrogen_chain(6), S.atom(10)]
    W, B = 70, 5

    def run(schedule):
        dev = [M.slater_jastrow(s, W) for s in systems]
        with M.Context(dev, seed=6, step=0.6, block_size=B, flags=M.FLAG_TRACE, record_sweeps=30,
                       precision=M.PREC_MIXED, lu_period=16) as ctx:
            for k, L in enumerate((4, 8, 16, 8)):
                ctx.set_walker_lanes(L, sys=k)
            ctx.set_tile_schedule(schedule, workgroups)
            if schedule == M.SCHEDULE_PERSISTENT:
                wg, _ = ctx.launch_info()
                assert wg == (workgroups or wg) and wg >= 1
            ctx.init_walkers()
            ctx.run(10)
            g = ctx.capture(5, sample=True)
            for _ in range(3):
                g.launch(1)
                ctx.synchronize()
                ctx.sweep(1)
            g.close()
            return [(ctx.get_configs(k), ctx.trace(28, 24, sys=k), ctx.stats(k)) for k in range(len(systems))]

    a = run(M.SCHEDULE_CLASSES)
    b = run(M.SCHEDULE_PERSISTENT)
    for k in range(len(systems)):
        assert np.array_equal(bits(a[k][0]), bits(b[k][0])), k
        assert np.array_equal(a[k][1][0], b[k][1][0]), k
        assert np.array_equal(bits(a[k][1][1]), bits(b[k][1][1])), k
        assert a[k][2] == b[k][2], k


@pytest.mark.parametrize("make,precision,proposal", [
    (lambda: S.atom(3), M.PREC_FP64, M.PROPOSAL_BOX),
    (lambda: S.atom(8), M.PREC_MIXED, M.PROPOSAL_BOX),
    (lambda: S.h2_lcao(), M.PREC_FP64, M.PROPOSAL_DIFFUSE),
])
def test_exponent_scale_parameter_parity_with_oracle(make, precision, proposal):
    # six parameters: the basis exponents are s zeta_m on both sides (same
    # products), so the trajectories match bit for bit, and the per-walker
    # sums of the six log-derivatives match the oracle's samples
    system = make()
    s0 = 1.15
    W, steps, B = 48, 24, 4
    side = 0.6 if proposal == M.PROPOSAL_BOX else 0.05
    p6 = list(system.jastrow) + [s0]
    dev = [M.slater_jastrow(system, W, params=p6)]
    with M.Context(dev, seed=2, step=side, block_size=B, flags=M.FLAG_TRACE, record_sweeps=steps,
                   precision=precision, lu_period=16, proposal=proposal) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        total = (steps // B) * B
        masks, el = ctx.trace(total, total - B)
        cfg = ctx.get_configs()
        sr = ctx.sr_sums()
        assert np.array_equal(ctx.get_params(), p6)
    wv = SJ.SjWave(system, params=p6)
    o = SJ.run(wv, seed=2, walkers=W, steps=steps, block_size=B, box_side=side, tau=side,
               precision=SJ.PREC_FP64 if precision == M.PREC_FP64 else SJ.PREC_MIXED, lu_period=16,
               n_threads=8, proposal=SJ.BOX if proposal == M.PROPOSAL_BOX else SJ.DIFFUSE)
    assert np.array_equal(masks, o["accept_mask"])
    assert np.array_equal(bits(cfg), bits(o["cfg"]))
    rel = np.abs(el - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= E_TOL
    dl = o["dlog"].reshape(-1, 6)
    assert sr["sum_o"].shape == (6,) and abs(sr["sum_o"][5]) > 0.0
    assert np.allclose(sr["sum_o"], dl.sum(axis=0), rtol=1e-9, atol=1e-9)
    assert np.allclose(sr["sum_oo"], dl.T @ dl, rtol=1e-9, atol=1e-9)
    assert np.allclose(sr["sum_oe"], dl.T @ o["e_local"].ravel(), rtol=1e-9, atol=1e-7)


def test_sho_steepest_descent_reaches_the_ground_state():
    # tests/sho_optimize.rs:117-141 on the device: a 3D Gaussian trial
    # function exp(-s r^2) (a = 1: s = 1) under T + r^2/2 with the drift-
    # diffusion proposal, steepest descent on the exponent from the energy
    # gradient 2(<E O> - <E><O>); the ground state s = 1/2 has E = 3/2 with
    # zero variance.  (The reference's step 1e-3 applies to its Q1-scaled
    # gradient; with the log-derivative a step of 0.2 converges in ~10
    # iterations.)  The last energy lies within its blocking error of 1.5.
    system = S.harmonic_oscillator(1.0, a=1.0)
    h = M.harmonic_hamiltonian(1.0)
    p = np.array(list(system.jastrow) + [1.0])
    dev = [M.slater_jastrow(system, 1024, params=p)]
    energies = []
    with M.Context(dev, seed=0, step=0.1, block_size=10, proposal=M.PROPOSAL_DIFFUSE) as ctx:
        ctx.set_hamiltonian(h)
        ctx.init_walkers()
        ctx.sweep(50)
        for it in range(20):
            ctx.reset_stats()
            ctx.run(100, 10)
            st = ctx.stats()
            mean = st["sum_e"] / st["n_samples"]
            nb = st["n_blocks"]
            err = np.sqrt(max(st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2, 0.0) / (nb - 1))
            energies.append((mean, err, p[5]))
            dp, g = M.parameter_update(ctx.sr_sums_raw(), M.OPT_SD, 0.2)
            assert np.all(dp[:5] == 0.0)  # no Jastrow term: only the exponent moves
            p = p + dp
            ctx.set_params(p)
    e0, err0, s_first = energies[0]
    e, err, s_last = energies[-1]
    assert e0 > 1.8 and s_first == 1.0  # E(s) = 3/2 (s + 1/(4 s)) = 1.875 at s = 1
    assert abs(s_last - 0.5) < 1e-6
    assert abs(e - 1.5) <= max(err, 1e-12), (e, err)


@pytest.mark.parametrize("make,params,optimizer,step,precision", [
    (lambda: S.helium_product(1.6, 0.4), [0.4, 0.0, 0.3, 0.0, 1.0, 1.05], M.OPT_SD, 0.05, M.PREC_FP64),
    (lambda: S.atom(3), None, M.OPT_SD, 0.01, M.PREC_MIXED),
    (lambda: S.atom(3), list(S.atom(3).jastrow) + [1.0], M.OPT_SR, 0.02, M.PREC_MIXED),
    (lambda: S.harmonic_oscillator(1.0, a=1.0), [0.0, 0.0, 0.0, 0.0, 0.0, 1.0], M.OPT_SD, 0.2, M.PREC_FP64),
])
def test_device_optimisation_loop_equals_host_loop(make, params, optimizer, step, precision):
    # mole_optimize on a Slater-Jastrow system (the step kernel rewrites the
    # device system table and, with the exponent scale, the exponents)
    # against the same iterations through the host (sums read back, the
    # update on the host, mole_set_params): the same parameters, sums and
    # walkers bit for bit
    system = make()
    p0 = np.array(system.jastrow if params is None else params, dtype=float)
    W, steps, B, iters = 256, 40, 10, 4
    h = M.harmonic_hamiltonian(1.0) if system.charges.sum() == 0 else None

    def context():
        c = M.Context([M.slater_jastrow(system, W, params=p0)], seed=9, step=0.5, block_size=B,
                      precision=precision, lu_period=16)
        if h is not None:
            c.set_hamiltonian(h)
        c.init_walkers()
        return c

    with context() as ctx:
        p = p0.copy()
        host_log = []
        for _ in range(iters):
            ctx.reset_stats()
            ctx.run(steps, B)
            st = ctx.stats()
            host_log.append((st["sum_e"], st["sum_block_mean"], st["n_accepted"]))
            dp, _ = M.parameter_update(ctx.sr_sums_raw(), optimizer, step, M.SR_TEXTBOOK, 1e-3)
            p = p + dp
            ctx.set_params(p)
        cfg_host = ctx.get_configs()
    with context() as ctx:
        p_dev, log = ctx.optimize(iters, steps, B, optimizer, step, M.SR_TEXTBOOK, 1e-3)
        assert np.array_equal(ctx.get_params(), p_dev)
        cfg_dev = ctx.get_configs()
    assert np.array_equal(bits(p_dev), bits(p))
    for it in range(iters):
        assert log[it, 0] == host_log[it][0] and log[it, 2] == host_log[it][1], it
        assert log[it, 4] == host_log[it][2] and log[it, 5] == 0.0
    assert np.array_equal(bits(cfg_dev), bits(cfg_host))
    assert not np.array_equal(p_dev, p0)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_compat_samples_match_oracle(precision):
    # MOLE_FLAG_SAMPLES on a Slater-Jastrow system: per-sample "Energy",
    # "Wavefunction value" (psi psi)/psi from the carried determinants and
    # "Parameter gradient" O psi_s, against the oracle's parity runner
    system = S.random_molecule(6, seed=5)
    W, steps, B = 70, 30, 10
    ns = (steps // B - 1) * B
    with M.Context([M.slater_jastrow(system, W)], seed=3, step=0.6, block_size=B, flags=M.FLAG_SAMPLES,
                   record_sweeps=steps, precision=precision, lu_period=16) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        e, p, g = ctx.samples(ns)
    o = run_oracle(system, W, steps, B, precision=SJ.PREC_FP64 if precision == M.PREC_FP64 else SJ.PREC_MIXED,
                   seed=3)
    rel = np.abs(e - o["e_local"]) / np.maximum(1.0, np.abs(o["e_local"]))
    assert np.max(rel) <= E_TOL
    assert np.all(np.isfinite(p)) and np.all(p != 0.0)
    assert np.array_equal(np.sign(p), np.sign(o["psi"]))
    assert np.max(np.abs(p - o["psi"]) / np.abs(o["psi"])) <= 1e-9
    want = o["dlog"][:, :, :5] * o["psi"][:, :, None]
    assert g.shape == (W, ns, 5)
    scale = np.abs(o["psi"])[:, :, None] * np.maximum(1.0, np.abs(o["dlog"][:, :, :5]))
    assert np.max(np.abs(g - want) / scale) <= 1e-9


def _configs_near_nuclei(system, n, seed):
    rng = np.random.default_rng(seed)
    N = system.n_electrons
    at = rng.integers(0, system.nuclei.shape[0], (n, N))
    return system.nuclei[at] + rng.normal(0.0, 0.8, (n, N, 3))


@pytest.mark.parametrize("make", [lambda: S.atom(4), lambda: S.atom(10), lambda: S.random_molecule(6, seed=5)])
@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_eval_matches_oracle(make, precision):
    # mole_eval on a Slater-Jastrow system (FP64 whatever the context's
    # precision): psi against the oracle's LU determinants and full Jastrow
    # sum, grad psi = psi drift, lap psi = psi kin with kin from E_L and the
    # potential, E_L and d psi/dp against the oracle's fresh-inverse local
    # energy; the context's own walkers are left as they were
    system = make()
    n, N = 37, system.n_electrons
    x = _configs_near_nuclei(system, n, 4)
    with M.Context([M.slater_jastrow(system, 8)], seed=1, step=0.6, block_size=2, precision=precision,
                   lu_period=16) as ctx:
        ctx.init_walkers()
        ctx.run(4)
        before = ctx.get_configs()
        out = ctx.eval(x)
        assert np.array_equal(bits(ctx.get_configs()), bits(before))
    wv = SJ.SjWave(system)
    R, Z = system.nuclei, system.charges.astype(float)
    v_nn = sum(Z[I] * Z[J] / np.linalg.norm(R[I] - R[J]) for I in range(len(Z)) for J in range(I + 1, len(Z)))
    for j in range(n):
        xj = x[j].ravel()
        v = SJ.value(wv, xj)
        e, dl = SJ.local_energy(wv, xj)
        assert abs(out["value"][j] - v) <= 1e-10 * abs(v), j
        assert abs(out["e_local"][j] - e) <= 1e-9 * max(1.0, abs(e)), j
        assert np.all(np.abs(out["pgrad"][j] - v * dl[:5]) <= 1e-9 * abs(v) * np.maximum(1.0, np.abs(dl[:5]))), j
        g = np.concatenate([SJ.drift(wv, xj, i) for i in range(N)])
        assert np.all(np.abs(out["grad"][j] - v * g) <= 1e-9 * abs(v) * np.maximum(1.0, np.abs(g))), j
        pos = x[j]
        V = -sum(Z[I] / np.linalg.norm(pos[i] - R[I]) for i in range(N) for I in range(len(Z)))
        V += sum(1.0 / np.linalg.norm(pos[i] - pos[k]) for i in range(N) for k in range(i + 1, N))
        kin = -2.0 * (e - V - v_nn)
        assert abs(out["lap"][j] / v - kin) <= 1e-8 * max(1.0, abs(kin)), j
