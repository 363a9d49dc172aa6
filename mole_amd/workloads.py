"""Drivers of the benchmark workloads beyond the default bench line.

atoms():   Li..Ne Slater-Jastrow atoms, 65,536 walkers each, Sherman-Morrison
           in FP32 with an FP64 Gauss-Jordan refresh every 64 moves, against
           an all-FP64 run on the same Philox streams.
hetero():  256 systems of 2..40 electrons (hydrogen chains and Poisson-disk
           molecules), walkers proportional to electron count, one launch per
           call, autotuned lanes per walker, sweeps replayed from a HIP graph.
learned(): the learned MH proposal trained on the device from box-sampler
           configurations, against the box proposal: tau_int, ESS/s, kappa.
"""
import time

import numpy as np

import mole_amd as M
from mole_amd import autotune, systems


def energy_and_error(st):
    n = st["n_samples"]
    mean = st["sum_e"] / n
    nb = st["n_blocks"]
    err = float(np.sqrt(max(st["sum_block_mean2"] / nb - (st["sum_block_mean"] / nb) ** 2, 0.0) / (nb - 1)))
    return mean, err


def run_atom(sj, walkers, sweeps, block, precision, lu_period, lanes, seed):
    dev = M.slater_jastrow(sj, walkers)
    with M.Context([dev], seed=seed, step=0.6, block_size=block, precision=precision,
                   lu_period=lu_period) as ctx:
        if lanes:
            ctx.set_walker_lanes(lanes)
        ctx.init_walkers()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(sweeps)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        kernel_ms, launches = ctx.last_timing()
        st = ctx.stats()
        lanes = ctx.walker_lanes()
    e, err = energy_and_error(st)
    moves = walkers * ((sweeps // block) * block) * sj.n_electrons
    acc = st["n_accepted"] / st["n_moves"]
    tflops = moves * sj_flop_per_move(sj, acc) / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else None
    return dict(energy=e, error=err, acceptance=acc,
                max_condition=st["max_condition"], n_singular=st.get("n_singular", 0),
                max_logdet_drift=st.get("max_logdet_drift", 0.0), seconds=wall, kernel_ms=kernel_ms,
                moves_per_s=moves / wall, lanes=lanes,
                roofline=dict(bound="fp64-valu", flop_per_move=sj_flop_per_move(sj, acc), achieved=tflops,
                              peak=78.6, unit="TFLOP/s", frac=None if tflops is None else tflops / 78.6))


def atoms(Zs=range(3, 11), walkers=65536, sweeps=1000, block=50, lu_period=64, seed=1, tune=True,
          cpu_sample=None, log=print):
    out = []
    for Z in Zs:
        sj = systems.atom(Z)
        L = 0
        if tune:
            table, _ = autotune.tune([sj], lambda s: walkers, candidates=(1, 2, 4, 8, 16),
                                     sweeps=4, precision=M.PREC_MIXED, lu_period=lu_period)
            L = table[sj.n_electrons]
        mixed = run_atom(sj, walkers, sweeps, block, M.PREC_MIXED, lu_period, L, seed)
        fp64 = run_atom(sj, walkers, sweeps, block, M.PREC_FP64, lu_period, L, seed)
        diff = mixed["energy"] - fp64["energy"]
        sig = float(np.hypot(mixed["error"], fp64["error"]))
        rec = dict(workload="atom", Z=Z, element=sj.name, n_up=sj.n_up, n_dn=sj.n_dn,
                   walkers=walkers, sweeps=sweeps, block=block, lu_period=lu_period,
                   mixed=mixed, fp64=fp64, mixed_minus_fp64=diff, sigma=sig,
                   within_1sigma=bool(abs(diff) <= sig),
                   fp32_trusted=bool(fp64["max_condition"] < 1e5))
        if cpu_sample:
            rec["cpu_baseline"] = cpu_sample(sj)
        log(rec)
        out.append(rec)
    return out


def hetero(n_systems=256, total_walkers=1 << 20, equil_sweeps=20, sweeps=100, chunk=10, block=10,
           precision=M.PREC_MIXED, lu_period=64, seed=3, tune=True, table=None, lds_budget=40 << 10,
           schedule=M.SCHEDULE_CLASSES, workgroups=0, tune_batch=True, log=print):
    """All systems in one context: per `chunk` sweeps one launch per LDS
    footprint class of the tiles (most expensive first within a class), or
    one persistent launch over every tile (schedule), replayed from a HIP
    graph.  tune: lanes per walker per electron count from single-system
    timings; tune_batch (with tune): the LDS budget, the schedule and the
    sweeps per launch from timings of the whole batch (overrides
    lds_budget, schedule and chunk)."""
    sys_list = hetero_systems(n_systems, seed)
    Ns = np.array([s.n_electrons for s in sys_list])
    W = np.maximum(1, np.floor(Ns / Ns.sum() * total_walkers)).astype(int)
    W[np.argmax(Ns)] += total_walkers - W.sum()
    timings = {}
    t_tune = 0.0
    if table is None:
        table = {}
        if tune:
            t0 = time.perf_counter()
            table, timings = autotune.tune(sys_list, lambda s: 16384, sweeps=4, precision=precision,
                                           lu_period=lu_period, lds_budget=None if tune_batch else lds_budget)
            t_tune = time.perf_counter() - t0
    dev = [M.slater_jastrow(s, int(w)) for s, w in zip(sys_list, W)]
    ctx = M.Context(dev, seed=seed, step=0.6, block_size=block, precision=precision, lu_period=lu_period)
    autotune.apply(ctx, sys_list, table)
    ctx.set_tile_schedule(schedule, workgroups)
    t0 = time.perf_counter()
    ctx.init_walkers()
    ctx.synchronize()
    t_init = time.perf_counter() - t0
    batch = None
    if tune and tune_batch and timings:
        t0 = time.perf_counter()
        best, trials = autotune.tune_batch(ctx, sys_list, timings)
        batch = dict(seconds=time.perf_counter() - t0, best={k: v for k, v in best.items() if k != "table"},
                     trials=trials)
        table, schedule, chunk = best["table"], best["schedule"], best["chunk"]
    geq = ctx.capture(chunk, sample=False)
    geq.launch(max(1, equil_sweeps // chunk))
    ctx.synchronize()
    ctx.reset_stats()
    g = ctx.capture(chunk, sample=True)
    reps = max(1, sweeps // chunk)
    t0 = time.perf_counter()
    g.launch(reps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    graph_ms, _ = ctx.last_timing()
    done = reps * chunk
    moves = int(np.sum(W * Ns)) * done
    n_wg, lds = ctx.launch_info()
    per_sys = []
    for k, s in enumerate(sys_list):
        st = ctx.stats(k)
        e, err = energy_and_error(st)
        per_sys.append(dict(name=s.name, N=int(Ns[k]), walkers=int(W[k]), lanes=ctx.walker_lanes(k),
                            energy=e, error=err, acceptance=st["n_accepted"] / st["n_moves"],
                            max_condition=st["max_condition"]))
    # the same sweeps launched eagerly, for the graph's overhead saving
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.sweep(chunk, sample=True)
    ctx.synchronize()
    wall_eager = time.perf_counter() - t0
    g.close()
    geq.close()
    ctx.close()
    flop = sweeps_flop(sys_list, W, done, per_sys)
    achieved = flop / (graph_ms * 1e-3) / 1e12 if graph_ms > 0 else None  # TFLOP/s over the graph replays
    rec = dict(workload="hetero", schedule="persistent" if schedule == M.SCHEDULE_PERSISTENT else "classes",
               systems=n_systems, walkers=int(W.sum()),
               electrons_min=int(Ns.min()), electrons_max=int(Ns.max()), sweeps=done, chunk=chunk,
               precision="mixed" if precision == M.PREC_MIXED else "fp64", lu_period=lu_period,
               seconds=wall, graph_device_ms=graph_ms, moves_per_s=moves / wall,
               eager_seconds=wall_eager, eager_moves_per_s=moves / wall_eager, init_seconds=t_init,
               workgroups_per_launch=n_wg, lds_per_workgroup=lds,
               lanes_table={int(k): int(v) for k, v in table.items()}, per_system=per_sys,
               tune_seconds=t_tune, autotune_batch=batch,
               roofline=dict(bound="fp64-valu", flop_per_sweep=flop / done, achieved=achieved, peak=78.6,
                             unit="TFLOP/s", frac=None if achieved is None else achieved / 78.6,
                             model="F(N, n, K, a) per electron-move, sj_flop_per_move"))
    log(rec)
    return rec


def sj_flop_per_move(system, acceptance):
    """Algorithmic FLOP of one electron-move of a Slater-Jastrow walker plus
    its share of the sampled sweep's local energy (N electrons, K basis
    functions, n = the moved electron's spin-block size, a = acceptance):
    F = 33 + 2K + 2nK + 2n + 14(N-1) + a(4n^2 + 2n)
        + [N(7K + 10nK + 8n + 36) + 21 N(N-1)/2] / N + 8
    (trial position and distance 33; basis values 2K and the new row of the
    orbital matrix 2nK; ratio 2n; Jastrow pairs 14(N-1); Sherman-Morrison
    4n^2 + 2n when accepted; the sampled sweep's gradients, laplacians and
    pair terms per electron; the FP64 re-inversion every 64 moves ~8)."""
    N = system.n_electrons
    K = len(system.zeta)
    F = 0.0
    for n in (system.n_up, system.n_dn):
        if n == 0:
            continue
        F += n * (33 + 2 * K + 2 * n * K + 2 * n + 14 * (N - 1) + acceptance * (4 * n * n + 2 * n)
                  + (N * (7 * K + 10 * n * K + 8 * n + 36) + 21 * N * (N - 1) / 2) / N + 8)
    return F / N


def sweeps_flop(sys_list, W, sweeps, per_sys):
    """Algorithmic FLOP of `sweeps` sampled sweeps of the batch."""
    return float(sum(w * s.n_electrons * sweeps * sj_flop_per_move(s, p["acceptance"])
                     for s, w, p in zip(sys_list, W, per_sys)))


def hetero_systems(n_systems=256, seed=3, max_electrons=40, basis_per_nucleus=2):
    """Hydrogen chains and Poisson-disk random molecules with 2..40 electrons."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_systems):
        N = int(rng.integers(2, max_electrons + 1))
        if k % 4 == 0:
            out.append(systems.hydrogen_chain(N, spacing=1.4, basis_per_nucleus=basis_per_nucleus, seed=seed + k))
        else:
            out.append(systems.random_molecule(N, seed=seed * 1000 + k, min_sep=1.2,
                                               basis_per_nucleus=basis_per_nucleus))
    return out


def learned(Z=10, walkers=65536, n_configs=2_000_000, collect_every=5, equil=1000, train_steps=24000,
            batch=16384, lr=1e-3, sweeps=400, block=50, box_sides=(0.1, 0.15, 0.2, 0.3, 0.45, 0.6, 0.9, 1.5),
            precision=M.PREC_FP64, lu_period=64, seed=1, objective=M.MLP_OBJ_LOG_ESJD, collect_side=0.3,
            ref_sweeps=20000, ref_block=500, log=print):
    """Learned proposal against the box proposal on one atom: configurations
    from the box sampler, Adam training on the device, then runs of each
    proposal from the same ensemble.  Efficiency per proposal: ESS/s =
    W S / (tau_int t) with tau_int of each walker's E_L series by the statfor
    estimator (scripts/statfor.rs:30-54) on the device, averaged over walkers,
    and kappa = 1 / (sigma(E)^2 t) with the blocking error; t is the device
    time of the S sampled sweeps.  Every run first equilibrates for `equil`
    sweeps of its own proposal.  Energies: the learned run against the box
    sides of the largest kappa (matched sigma(E): kappa is 1/(sigma^2 t))
    and of the largest ESS/s, in units of the combined error (the larger of
    the blocking and the tau_int-based error of each), and against a long
    reference run of the box side with the shortest tau_int (5 x equil
    sweeps of equilibration, then ref_sweeps sampled sweeps in blocks of
    ref_block, far longer than tau_int, so its blocking error holds where
    the short box chains of small sides have not mixed)."""
    sj = systems.atom(Z)
    N = sj.n_electrons

    def context(proposal, step, flags=0, record=0):
        return M.Context([M.slater_jastrow(sj, walkers)], seed=seed, step=step, block_size=block,
                         precision=precision, lu_period=lu_period, proposal=proposal, flags=flags,
                         record_sweeps=record)

    mlp = M.Mlp(seed=2)
    # 1. training configurations from the plain sampler; its final ensemble
    #    is also the common start of every measured run (a slowly mixing
    #    proposal started from U(-1,1) would otherwise carry its burn-in)
    with context(M.PROPOSAL_BOX, collect_side) as ctx:
        ctx.init_walkers()
        ctx.sweep(equil)
        n_snap = max(1, -(-n_configs // walkers))
        t0 = time.perf_counter()
        mlp.collect(ctx, n_snap, collect_every)
        ctx.synchronize()
        t_collect = time.perf_counter() - t0
        start = ctx.get_configs()
        # 2. training on the device
        t0 = time.perf_counter()
        obj = mlp.train(ctx, train_steps, batch=batch, lr=lr, seed=seed + 10, objective=objective)
        t_train = time.perf_counter() - t0
    n_data = mlp.data_size()
    box_lanes = autotune.tune([sj], lambda s: walkers, candidates=(4, 8, 16), sweeps=2, precision=precision,
                              lu_period=lu_period)[0]

    def measure(proposal, step):
        with context(proposal, step, M.FLAG_TRACE, sweeps + block) as ctx:
            if proposal == M.PROPOSAL_LEARNED:
                mlp.attach(ctx)
            else:
                autotune.apply(ctx, [sj], box_lanes)
            ctx.set_configs(start)
            ctx.sweep(equil)  # the proposal's own equilibration (small box sides mix slowly)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.run(sweeps + block)
            ctx.synchronize()
            wall = time.perf_counter() - t0
            dev_ms, _ = ctx.last_timing()
            st = ctx.stats()
            tau = ctx.tau_int(sweeps)  # on the device, from the trace
            lanes = ctx.walker_lanes()
        t = dev_ms * 1e-3 * sweeps / (sweeps + block)  # device time of the sampled sweeps
        e, err = energy_and_error(st)
        tau_mean = float(np.mean(tau))
        # error from the variance and tau_int (blocks shorter than the
        # correlation time understate the blocking error)
        var = max(st["sum_e2"] / st["n_samples"] - e * e, 0.0)
        err_tau = float(np.sqrt(var * tau_mean / st["n_samples"]))
        return dict(proposal={M.PROPOSAL_BOX: "box", M.PROPOSAL_LEARNED: "learned"}[proposal], step=step,
                    lanes=lanes, energy=e, error=err, error_tau=err_tau, acceptance=st["n_accepted"] / st["n_moves"],
                    tau_int_mean=tau_mean, tau_int_median=float(np.median(tau)), seconds=t, wall=wall,
                    moves_per_s=walkers * sweeps * N / t, ess_per_s=walkers * sweeps / (tau_mean * t),
                    kappa=1.0 / (err * err * t), max_condition=st["max_condition"])

    runs = [measure(M.PROPOSAL_BOX, s) for s in box_sides]
    for r in runs:
        log(dict(workload="learned-run", Z=Z, **r))
    lrn = measure(M.PROPOSAL_LEARNED, 0.0)
    log(dict(workload="learned-run", Z=Z, **lrn))
    mlp.close()
    best_ess = max(runs, key=lambda r: r["ess_per_s"])
    best_kappa = max(runs, key=lambda r: r["kappa"])
    best_tau = min(runs, key=lambda r: r["tau_int_mean"])
    ref = None
    if ref_sweeps:
        with M.Context([M.slater_jastrow(sj, walkers)], seed=seed + 1, step=best_tau["step"], block_size=ref_block,
                       precision=precision, lu_period=lu_period, proposal=M.PROPOSAL_BOX) as ctx:
            autotune.apply(ctx, [sj], box_lanes)
            ctx.set_configs(start)
            ctx.sweep(5 * equil)
            ctx.run(ref_sweeps + ref_block)
            e_ref, err_ref = energy_and_error(ctx.stats())
        ref = dict(step=best_tau["step"], sweeps=ref_sweeps, block=ref_block, energy=e_ref, error=err_ref)
        log(dict(workload="learned-reference", Z=Z, **ref))
    diff = lrn["energy"] - best_kappa["energy"]
    sig = lambda r: max(r["error"], r["error_tau"])
    diff_ess = lrn["energy"] - best_ess["energy"]
    rec = dict(workload="learned", Z=Z, element=sj.name, walkers=walkers, sweeps=sweeps, block=block,
               precision="mixed" if precision == M.PREC_MIXED else "fp64",
               training=dict(configs=n_data, steps=train_steps, batch=batch, lr=lr,
                             objective={M.MLP_OBJ_ESJD: "esjd", M.MLP_OBJ_ESJD_RELATIVE: "esjd-relative",
                                        M.MLP_OBJ_LOG_ESJD: "log-esjd"}[objective],
                             objective_first=float(np.mean(obj[:20])), objective_last=float(np.mean(obj[-20:])),
                             collect_seconds=t_collect, train_seconds=t_train),
               box=runs, learned=lrn,
               best_box_side_ess=best_ess["step"], best_box_side_kappa=best_kappa["step"],
               ess_ratio_vs_best_box=lrn["ess_per_s"] / best_ess["ess_per_s"],
               kappa_ratio_vs_best_box=lrn["kappa"] / best_kappa["kappa"],
               best_box_side_tau=best_tau["step"],
               tau_ratio_vs_best_box=best_tau["tau_int_mean"] / lrn["tau_int_mean"],
               energy_diff_vs_box=diff,
               energy_diff_sigma=diff / float(np.hypot(sig(lrn), sig(best_kappa))),
               energy_diff_vs_best_ess_box=diff_ess,
               energy_diff_sigma_vs_best_ess_box=diff_ess / float(np.hypot(sig(lrn), sig(best_ess))))
    if ref is not None:
        rec.update(reference=ref, energy_diff_sigma_vs_reference=(lrn["energy"] - ref["energy"]) /
                   float(np.hypot(sig(lrn), ref["error"])))
    log(rec)
    return rec


DMC_EXACT = {"He product": -2.903724, "H2 LCAO": -1.174476}


def dmc_systems():
    """Nodeless trial functions whose local energy stays bounded at the
    nuclei: He with exp(-2 r) orbitals (the electron-nucleus cusp) and a
    Pade e-e factor (the electron-electron cusp), and H2 in the sigma_g LCAO
    orbital; fixed-node DMC is exact for both ground states.  (The random
    orbitals of the atoms workload break the cusps and have nodes: their
    local energies diverge, weights overflow, and SRBrancher's total
    integer weight vanishes -- the reference panics there too.)"""
    return [systems.helium_product(2.0, 0.5, 0.3), systems.h2_lcao()]


def dmc(sys_list=None, walkers=65536, iters=2000, block=100, eq=2, tau=0.01, brancher=M.BRANCH_SR,
        precision=M.PREC_MIXED, lu_period=64, seed=5, cpu_sample=None, log=print):
    """DmcRunner::diffuse on the device, one system per context: a VMC
    estimate as the trial energy, then `iters` DMC steps (each a captured
    graph of the sweep launch, the reductions, branching and bookkeeping).
    Rate = walkers x steps x electrons / device time of the diffuse call."""
    out = []
    for sj in (sys_list or dmc_systems()):
        dev = M.slater_jastrow(sj, walkers)
        table, _ = autotune.tune([sj], lambda s: walkers, candidates=(1, 2, 4, 8), sweeps=4, precision=precision,
                                 lu_period=lu_period)
        with M.Context([dev], seed=seed, step=tau, block_size=50, precision=precision, lu_period=lu_period,
                       proposal=M.PROPOSAL_DIFFUSE) as ctx:
            autotune.apply(ctx, [sj], table)
            lanes = ctx.walker_lanes()
            ctx.init_walkers()
            ctx.sweep(100)
            ctx.reset_stats()
            ctx.run(200, 50)
            e_vmc, err_vmc = energy_and_error(ctx.stats())
            ctx.dmc_init(e_vmc, brancher)
            ctx.dmc_diffuse(tau, 20, 10, 0)  # warm-up (not timed)
            ctx.reset_stats()
            t0 = time.perf_counter()
            e, err = ctx.dmc_diffuse(tau, iters, block, eq)
            wall = time.perf_counter() - t0
            dev_ms, _ = ctx.last_timing()
            w, _ = ctx.dmc_read_walkers()
            st = ctx.stats()
        moves = walkers * iters * sj.n_electrons
        rec = dict(workload="dmc", system=sj.name, N=sj.n_electrons, walkers=walkers, steps=iters, block=block,
                   eq_blocks=eq, tau=tau, brancher="SR" if brancher == M.BRANCH_SR else "simple",
                   precision="mixed" if precision == M.PREC_MIXED else "fp64",
                   vmc_energy=e_vmc, vmc_error=err_vmc,
                   dmc_energy=float(e[-1]) if len(e) else None, dmc_error=float(err[-1]) if len(err) else None,
                   exact=DMC_EXACT.get(sj.name),
                   lanes=lanes, mean_weight=float(np.mean(w)), acceptance=st["n_accepted"] / max(1, st["n_moves"]),
                   device_seconds=dev_ms * 1e-3, wall_seconds=wall, moves_per_s=moves / (dev_ms * 1e-3),
                   ms_per_step=dev_ms / iters)
        if cpu_sample:
            rec["cpu_baseline"] = cpu_sample(sj, e_vmc, tau, brancher)
        log(rec)
        out.append(rec)
    return out
