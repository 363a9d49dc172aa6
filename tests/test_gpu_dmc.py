"""Diffusion Monte Carlo on the device (mole_dmc_*) against the oracle's
DmcRunner restatement (DmcRunner::diffuse, src/dmc/src/dmc.rs:75-202;
SRBrancher / SimpleBranching, src/dmc/src/branching.rs:7-91) on the same
Philox streams: step energies, block energies, the reference energy, walker
weights, local energies and configurations bit for bit (the oracle sums over
walkers in the device's fixed order; the block errors, formed on the host
from the variances, to 1e-12).  Physics anchors: zero-variance hydrogen, the He
ground state; the panic path of SRBrancher."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import systems as S
from oracle import sj_ffi as SJ

pytestmark = pytest.mark.gpu

HE_EXACT = -2.903724


def run_device(sys_list, W, *, tau, time_step, iters, block, eq, eref, brancher, seed=3,
               precision=M.PREC_FP64, lu=0, lanes=None, hamiltonian=M.H_ELECTRONIC):
    dev = [M.slater_jastrow(s, W, hamiltonian=hamiltonian) for s in sys_list]
    with M.Context(dev, seed=seed, step=tau, block_size=10, precision=precision, lu_period=lu,
                   proposal=M.PROPOSAL_DIFFUSE) as ctx:
        if lanes is not None:
            for k in range(len(sys_list)):
                ctx.set_walker_lanes(lanes, sys=k)
        ctx.dmc_init(eref, brancher)
        ctx.dmc_diffuse(time_step, iters, block, eq)
        out = []
        for k in range(len(sys_list)):
            e, err, er = ctx.dmc_read(k)
            w, el = ctx.dmc_read_walkers(k)
            out.append(dict(e=e, err=err, eref=er, steps=ctx.dmc_read_steps(k), w=w, el=el,
                            cfg=ctx.get_configs(sys=k)))
    return out


def run_oracle(system, W, *, tau, time_step, iters, block, eq, eref, brancher, seed=3, sys=0,
               precision=SJ.PREC_FP64, lu=0, hamiltonian=0):
    wv = SJ.SjWave(system, hamiltonian=hamiltonian)
    d = SJ.Dmc(wv, walkers=W, seed=seed, tau=tau, reference_energy=eref, brancher=brancher, precision=precision,
               lu_period=lu, sys=sys)
    e, err, se = d.diffuse(time_step, iters, block, eq)
    return dict(e=e, err=err, eref=d.reference_energy, steps=se, w=d.weight.copy(), el=d.e_local.copy(),
                cfg=d.cfg.copy())


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def check_parity(d, o):
    # the local energies are bit-identical to the oracle's and the oracle sums
    # over walkers in the device's order, so everything agrees bit for bit
    assert np.array_equal(bits(d["steps"]), bits(o["steps"]))
    assert np.array_equal(bits(d["e"]), bits(o["e"]))
    assert np.allclose(d["err"], o["err"], rtol=1e-12, atol=1e-15)
    assert d["eref"] == o["eref"]
    assert np.array_equal(bits(d["w"]), bits(o["w"]))
    assert np.array_equal(bits(d["el"]), bits(o["el"]))
    assert np.array_equal(bits(d["cfg"].reshape(o["cfg"].shape)), bits(o["cfg"]))


@pytest.mark.parametrize("brancher", [M.BRANCH_SR, M.BRANCH_SIMPLE])
@pytest.mark.parametrize("system,eref,lu", [(S.helium_product(1.6875, 0.5, 0.3), -2.88, 0),
                                            (S.atom(3), -7.0, 8)])
def test_dmc_parity_with_oracle(brancher, system, eref, lu):
    kw = dict(tau=0.02, time_step=0.02, iters=30, block=5, eq=1, eref=eref, brancher=brancher, lu=lu)
    print(f"dmc parity {system.name} brancher={brancher}", flush=True)
    d = run_device([system], 96, **kw)[0]
    o = run_oracle(system, 96, **kw)
    check_parity(d, o)


@pytest.mark.parametrize("brancher", [M.BRANCH_SR, M.BRANCH_SIMPLE])
def test_dmc_mixed_precision_parity_with_oracle(brancher):
    # FP32 Sherman-Morrison inverses (kept as FP32 values by the oracle too),
    # FP64 local energies and weights
    kw = dict(tau=0.02, time_step=0.02, iters=30, block=5, eq=1, eref=-7.0, brancher=brancher, lu=8)
    d = run_device([S.atom(3)], 96, precision=M.PREC_MIXED, **kw)[0]
    o = run_oracle(S.atom(3), 96, precision=SJ.PREC_MIXED, **kw)
    check_parity(d, o)


@pytest.mark.parametrize("W,brancher,iters,block", [
    (9000, M.BRANCH_SR, 12, 4),        # 3 reduction partials, 9 scan blocks (offsets scanned in the selection)
    (9000, M.BRANCH_SIMPLE, 12, 4),    # list scans over up to 27,000 entries
    (263000, M.BRANCH_SR, 4, 2),       # 257 scan blocks: the separate block-offset scan
])
def test_dmc_parity_multi_block(W, brancher, iters, block):
    # walker counts that span several reduction and scan workgroups, ragged
    kw = dict(tau=0.02, time_step=0.02, iters=iters, block=block, eq=1, eref=-2.88, brancher=brancher)
    print(f"dmc multi-block parity W={W} brancher={brancher}", flush=True)
    d = run_device([S.helium_product(1.6875, 0.5, 0.3)], W, **kw)[0]
    o = run_oracle(S.helium_product(1.6875, 0.5, 0.3), W, **kw)
    check_parity(d, o)


@pytest.mark.parametrize("k,eref", [(0, -2.87), (1, -1.14)])
def test_dmc_parity_at_workload_shape(k, eref):
    # the DMC workload's shape: its trial functions (He exp(-2r) orbitals with
    # a Pade factor, H2 sigma_g LCAO), 65,536 walkers, 2 lanes per walker,
    # mixed precision with an LU refresh every 64 moves, SRBrancher, tau 0.01;
    # a few steps bit for bit against the oracle
    from mole_amd import workloads as WL
    system = WL.dmc_systems()[k]
    kw = dict(tau=0.01, time_step=0.01, iters=6, block=2, eq=1, eref=eref, brancher=M.BRANCH_SR, lu=64, seed=5)
    print(f"dmc workload shape {system.name} W=65536", flush=True)
    d = run_device([system], 65536, precision=M.PREC_MIXED, lanes=2, **kw)[0]
    o = run_oracle(system, 65536, precision=SJ.PREC_MIXED, **kw)
    check_parity(d, o)


def test_dmc_lane_independence():
    kw = dict(tau=0.02, time_step=0.02, iters=20, block=5, eq=1, eref=-7.0, brancher=M.BRANCH_SR, lu=8)
    a = run_device([S.atom(3)], 64, lanes=1, **kw)[0]
    b = run_device([S.atom(3)], 64, lanes=8, **kw)[0]
    assert np.array_equal(a["steps"], b["steps"]) and np.array_equal(a["w"], b["w"])
    assert np.array_equal(a["cfg"], b["cfg"])


def test_dmc_two_systems_match_their_own_oracle_runs():
    sys_list = [S.helium_product(1.6875, 0.5, 0.3), S.atom(4)]
    kw = dict(tau=0.02, time_step=0.02, iters=20, block=4, eq=1, brancher=M.BRANCH_SR, lu=8)
    d = run_device(sys_list, 80, eref=[-2.88, -14.0], **kw)
    for k, (s, er) in enumerate(zip(sys_list, [-2.88, -14.0])):
        o = run_oracle(s, 80, eref=er, sys=k, **kw)
        check_parity(d[k], o)


def test_dmc_zero_variance_hydrogen():
    d = run_device([S.hydrogenic(1)], 128, tau=0.05, time_step=0.05, iters=40, block=10, eq=1, eref=-0.5,
                   brancher=M.BRANCH_SR)[0]
    assert np.all(np.abs(d["steps"] + 0.5) < 1e-12)
    assert np.all(d["w"] == 1.0) and abs(d["eref"] + 0.5) < 1e-12


def test_dmc_zero_variance_helium_ionic():
    d = run_device([S.helium_product(2.0, 0.0)], 128, tau=0.02, time_step=0.02, iters=30, block=10, eq=1,
                   eref=-4.0, brancher=M.BRANCH_SIMPLE, hamiltonian=M.H_IONIC)[0]
    assert np.all(np.abs(d["steps"] + 4.0) < 1e-11)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_dmc_helium_ground_state(precision):
    d = run_device([S.helium_product(1.6875, 0.5, 0.3)], 4096, tau=0.01, time_step=0.01, iters=3000, block=100,
                   eq=5, eref=-2.88, brancher=M.BRANCH_SR, precision=precision)[0]
    e, err = d["e"][-1], d["err"][-1]
    print(f"He DMC precision={precision}: {e:.5f} +- {err:.5f}", flush=True)
    assert abs(e - HE_EXACT) < 3 * err + 0.004


def test_dmc_vanishing_weights_is_an_error():
    # every weight underflows to 0 in the first step: the reference's
    # WeightedChoice panics (total weight 0); the device reports it
    with pytest.raises(M.MoleError) as ei:
        run_device([S.helium_product(1.6875, 0.5, 0.3)], 64, tau=0.02, time_step=1e4, iters=2, block=1, eq=0,
                   eref=-1e3, brancher=M.BRANCH_SR)
    assert ei.value.status == M.ERR_INVALID_ARGUMENT
    wv = SJ.SjWave(S.helium_product(1.6875, 0.5, 0.3))
    o = SJ.Dmc(wv, walkers=64, seed=3, tau=0.02, reference_energy=-1e3)
    with pytest.raises(ValueError):
        o.diffuse(1e4, 2, 1, 0)


def test_dmc_requires_init_and_diffusion():
    with M.Context([M.slater_jastrow(S.atom(3), 16)], seed=1, step=0.5, block_size=10) as ctx:
        with pytest.raises(M.MoleError) as ei:
            ctx.dmc_init(-7.0)
        assert ei.value.status == M.ERR_UNSUPPORTED
    with M.Context([M.slater_jastrow(S.atom(3), 16)], seed=1, step=0.02, block_size=10,
                   proposal=M.PROPOSAL_DIFFUSE) as ctx:
        with pytest.raises(M.MoleError):
            ctx.dmc_diffuse(0.02, 10, 5, 1)


def test_dmc_graph_replays_equal_eager_steps():
    # a diffuse call of 20 steps replays one captured step 20 times; 20 calls
    # of one step each run eagerly.  With num_eq_blocks above the block count
    # E_ref never changes, so both schedules are the same trajectory, bit
    # for bit, replay after replay
    sj = S.atom(3)
    runs = []
    for mode in ("graph", "eager"):
        with M.Context([M.slater_jastrow(sj, 64)], seed=2, step=0.02, block_size=10, lu_period=8,
                       proposal=M.PROPOSAL_DIFFUSE) as ctx:
            ctx.dmc_init(-7.0, M.BRANCH_SIMPLE)
            if mode == "graph":
                ctx.dmc_diffuse(0.02, 20, 1, 100)
                steps = ctx.dmc_read_steps()
            else:
                steps = []
                for _ in range(20):
                    ctx.dmc_diffuse(0.02, 1, 1, 100)
                    steps.append(ctx.dmc_read_steps()[0])
                steps = np.array(steps)
            w, el = ctx.dmc_read_walkers()
            runs.append((steps, w, el, ctx.get_configs()))
    (gs, gw, ge, gc), (es, ew, ee, ec) = runs
    assert gs.size == es.size == 20
    assert np.array_equal(gs, es) and np.array_equal(gw, ew) and np.array_equal(ge, ee)
    assert np.array_equal(gc, ec)
