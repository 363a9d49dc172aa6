"""Walker-moves/s of the VMC hot path on one MI355X, and effective samples/s
of the learned proposal against the plain one.

Workload: H2 molecule (nuclei at z = +-0.7 bohr) with the LCAO sigma_g^2 x
Pade-Jastrow wavefunction, optimised by stochastic reconfiguration: every
step is one SR iteration = 16,384 walkers x 200 box-Metropolis sweeps
(block 20, first block discarded), local energy + parameter log-derivatives
sampled after every sweep, device-reduced SR sums, 3x3 solve, step 0.05,
diagonal shift 1e-3, FP64 throughout.  Walkers persist across iterations.
The default run times the full 50 iterations.

The line's "learned" record is the metric's second half on the Ne atom
(10 electrons, 65,536 walkers): the MLP proposal trained on the device on
2M configurations of the box sampler, then effective samples/s and
kappa = 1/(sigma(E)^2 t) against the best box side, tau_int on the device.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

WALKERS = 16384
SWEEPS = 200
BLOCK = 20
BOX = 0.5
SR_STEP = 0.05
SR_EPS = 1e-3
PARAMS0 = (1.0, 0.5, 0.3)  # zeta, b, c
N_ELEC = 2

# Algorithmic work of the sweep kernel (+ - * = 1, fma = 2, / sqrt exp = 1):
# per electron-move 52 (3 Range maps 6, Standard 1, x' 3, |x'-A| and |x'-B|
# 18, two exp and the sum 5, r12' 9, Pade + exp 5, psi' 2, ratio 3); per
# sampled sweep 217 (E_L and d ln psi/dp 192, accumulators 25).
FLOP_MOVE = 52
FLOP_SAMPLE = 217
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector


def flops_per_launch(walkers, sweeps, sampled):
    return walkers * (sweeps * N_ELEC * FLOP_MOVE + sampled * FLOP_SAMPLE)


VARIANTS = {"auto": 0, "lane": 1, "ws": 2, "quad": 3}

# HBM bytes per sweep launch of the timed optimisation loop, as measured by
# rocprofv3 (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a plain copy
# kernel of known size reads) and recorded with the command that made it in
# TRAFFIC_SOURCE; a counter run cannot share this process, so the line carries
# that measurement.  The launches start the accumulators from zero, so the
# algorithmic state traffic is positions in and out plus accumulators out:
# 3.93e6 B.
TRAFFIC_BYTES_PER_LAUNCH = (680.3 * 2.0 + 3328.0) * 1024
TRAFFIC_SOURCE = "profiles/sweep2_ws_kernel_h2_sr_16384w.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"
METRIC = "walker-moves/s incl. E_L; effective samples/s learned vs plain proposal at matched sigma(E)"


def host_cpus():
    """The host's logical CPUs, this process's share of them (affinity mask
    and cgroup quota) and the CPU model; the baseline runs one thread per CPU
    of the share, at most 16 (a GPU's share of its box)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = aff if quota is None else min(aff, max(1, int(quota)))
    return dict(nproc=nproc, affinity=aff, cgroup_quota=quota, share=share, model=model, threads=min(share, 16))


def gpu_run(iters, warmup, seed=0, variant="auto", host_loop=False):
    """The SR optimisation on the device.  Default: mole_optimize, the whole
    loop (sweeps, reductions, SR solve, parameter update) enqueued without a
    host round trip.  host_loop: one C-ABI round trip per iteration (read the
    sums, solve on the host, push the parameters) -- the same trajectory bit
    for bit.  Afterwards three more iterations through the host path time the
    sweep kernel alone with HIP events on the context's stream (the roofline's
    kernel time)."""
    import mole_amd as M

    def one_iteration(ctx, p):
        ctx.reset_stats()
        ctx.run(SWEEPS, BLOCK)
        raw = ctx.sr_sums_raw()
        st = ctx.stats()
        dp, _ = M.parameter_update(raw, M.OPT_SR, SR_STEP, M.SR_TEXTBOOK, SR_EPS)
        p = p + dp
        ctx.set_params(p)
        ms, launches = ctx.last_timing()
        return p, st, ms, launches

    def stats_of(st):
        mean = st["sum_e"] / st["n_samples"]
        nb = st["n_blocks"]
        err = np.sqrt(max(st["sum_block_mean2"] / nb - mean**2, 0.0) / (nb - 1))
        return mean, err, st["n_accepted"] / st["n_moves"]

    def optimize(ctx, n):
        return ctx.optimize(n, SWEEPS, BLOCK, M.OPT_SR, SR_STEP, M.SR_TEXTBOOK, SR_EPS)

    # the timed context is set up first; the warmup runs on its own context
    # (the timed run starts from the initial parameters) right before the
    # timed loop, so the device is not idle between the two
    ctx = M.Context([M.h2(WALKERS, *PARAMS0)], seed=seed, step=BOX, block_size=BLOCK)
    ctx.set_launch_config(VARIANTS[variant])
    ctx.init_walkers()
    ctx.synchronize()
    w = None
    if warmup:  # closed after the timed loop: freeing it would pause the device
        w = M.Context([M.h2(WALKERS, *PARAMS0)], seed=seed + 1, step=BOX, block_size=BLOCK)
        w.set_launch_config(VARIANTS[variant])
        w.init_walkers()
        if host_loop:
            p = np.array(PARAMS0)
            for _ in range(warmup):
                p, *_ = one_iteration(w, p)
        else:
            optimize(w, warmup)
        w.synchronize()
    energies = []
    launches_total = 0
    t0 = time.perf_counter()
    if host_loop:
        p = np.array(PARAMS0)
        for _ in range(iters):
            p, st, ms, launches = one_iteration(ctx, p)
            launches_total += launches
            energies.append(stats_of(st))
    else:
        p, log = optimize(ctx, iters)
    ctx.synchronize()
    t1 = time.perf_counter()
    if w is not None:
        w.close()
    if not host_loop:
        _, launches_total = ctx.last_timing()
        nb = WALKERS * (SWEEPS // BLOCK - 1)
        for row in log:
            mean = row[0] / row[6]
            err = np.sqrt(max(row[3] / nb - mean**2, 0.0) / (nb - 1))
            energies.append((mean, err, row[4] / (WALKERS * SWEEPS * N_ELEC)))
    # the sweep kernel alone, timed live on the context's stream
    kernel_ms = []
    q = p.copy()
    for _ in range(3):
        q, _, ms, n_l = one_iteration(ctx, q)
        kernel_ms.append(ms / n_l)
    ctx.close()
    return dict(seconds=t1 - t0, kernel_ms=kernel_ms, launches=launches_total, energies=energies,
                params=[float(v) for v in p])


def cpu_baseline(threads, budget_s=20.0, host=None):
    """The oracle's reference-faithful VMC (two full psi evaluations per move,
    generic act_on/psi local energy, per-sample observables, SR from the stored
    samples) on the host, over whole SR iterations of the same workload."""
    import ctypes as C

    from oracle import oracle_ffi as O

    L = O.lib()

    def run(iters):
        wf = O.h2(*PARAMS0)
        op = O.OptParams(optimizer=O.OPT_SR, step=SR_STEP, sr_mode=O.SR_TEXTBOOK, sr_eps=SR_EPS,
                         o_convention=O.O_LOGDERIV, persistent=1)
        bufs = [np.zeros(iters * 3) for _ in range(5)]
        t0 = time.perf_counter()
        rc = L.orc_vmc_optimize(C.byref(wf), 0, BOX, iters, WALKERS * SWEEPS, BLOCK, WALKERS,
                                C.byref(op), threads, *[O.dptr(b) for b in bufs])
        dt = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle VMC failed")
        return dt

    t1 = run(1)
    iters = 1
    if t1 < budget_s / 2:
        iters = int(min(50, max(1, budget_s // t1)))
        if iters > 1:
            t1 = run(iters)
    moves = iters * WALKERS * SWEEPS * N_ELEC
    host = host or {}
    return dict(value=moves / t1, unit="walker-moves/s", cores=threads, kind="port",
                sample=f"{iters} SR iteration(s) of the same workload ({WALKERS} walkers x {SWEEPS} "
                       f"sweeps each) on the C oracle in reference-faithful mode (two full psi evaluations "
                       f"per move, the reference's generic local energy; the reference is Rust and cannot be "
                       f"built here), {threads} threads, {t1:.1f} s",
                nproc=host.get("nproc"), cpu_share=host.get("share"), cpu_model=host.get("model"),
                build="gcc -O3 -ffp-contract=off -fno-fast-math")


def helium_record(threads=None, cpu=True):
    """The plain-VMC helium singlet run in full: exp(-zeta (r1 + r2)) with the
    Pade e-e factor (zeta 1.6875, b 0.5, c 0.3), box moves of side 0.5,
    walkers uniform in [-1, 1]^3 per electron, 4,096 walkers x 10,000 sweeps,
    FP64, Philox seed 0 (blocks of 100 sweeps): E +- sigma and acceptance on
    the device, and the same chains on the oracle's reference-faithful runner
    (two full psi evaluations per move) on the host cores, whose energy must
    equal the device's (same streams, same decisions)."""
    import mole_amd as M

    walkers, sweeps, block = 4096, 10000, 100
    with M.Context([M.helium(walkers)], seed=0, step=0.5, block_size=block) as ctx:
        ctx.init_walkers()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(sweeps, block)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        dev_ms, launches = ctx.last_timing()
        st = ctx.stats()
    moves = walkers * sweeps * 2

    def energy(sum_e, n, sum_bm, sum_bm2, nb):
        mean = sum_e / n
        return mean, float(np.sqrt(max(sum_bm2 / nb - (sum_bm / nb) ** 2, 0.0) / (nb - 1)))

    e, err = energy(st["sum_e"], st["n_samples"], st["sum_block_mean"], st["sum_block_mean2"], st["n_blocks"])
    rec = dict(workload="He singlet plain VMC", walkers=walkers, sweeps=sweeps, block=block, box_side=0.5,
               params=[1.6875, 0.5, 0.3], energy=e, error=err, acceptance=st["n_accepted"] / st["n_moves"],
               device_seconds=dev_ms * 1e-3, wall_seconds=wall, launches=launches,
               moves_per_s=moves / (dev_ms * 1e-3), moves_per_s_wall=moves / wall)
    if cpu:
        from oracle import oracle_ffi as O

        threads = threads or host_cpus()["threads"]
        t0 = time.perf_counter()
        r = O.run_box(O.helium(), seed=0, walkers=walkers, steps=sweeps, block_size=block, box_side=0.5,
                      n_threads=threads, record=False)
        dt = time.perf_counter() - t0
        ec, errc = energy(r["sum_e"], r["n_samples"], r["sum_bm"], r["sum_bm2"], r["n_blocks"])
        rec.update(cpu=dict(seconds=dt, threads=threads, moves_per_s=moves / dt, energy=ec, error=errc,
                            acceptance=r["n_accepted"] / r["n_moves"], kind="port",
                            runner="oracle reference-faithful (two full psi evaluations per move)"),
                   gpu_vs_cpu=(moves / (dev_ms * 1e-3)) / (moves / dt), energy_gpu_minus_cpu=e - ec,
                   accepted_gpu_minus_cpu=int(st["n_accepted"]) - int(r["n_accepted"]))
    return rec


def cpu_atom_sample(threads, walkers=1024, sweeps=40, block=10):
    """Reference-faithful oracle (two full LU evaluations per move, fresh-inverse
    local energy) on a bounded sample of an atom's workload."""
    from oracle import sj_ffi as SJ

    def sample(sj):
        wv = SJ.SjWave(sj)
        t0 = time.perf_counter()
        SJ.run(wv, seed=1, walkers=walkers, steps=sweeps, block_size=block, box_side=0.6,
               faithful=True, n_threads=threads, record=False)
        dt = time.perf_counter() - t0
        moves = walkers * sweeps * sj.n_electrons
        return dict(value=moves / dt, unit="walker-moves/s", cores=threads, kind="port",
                    sample=f"{walkers} walkers x {sweeps} sweeps, oracle reference-faithful mode, {dt:.2f} s")
    return sample


def atoms_main(args):
    from mole_amd import workloads

    threads = args.cpu_threads or host_cpus()["threads"]
    cpu = None if args.no_cpu_baseline else cpu_atom_sample(threads)
    workloads.atoms(walkers=args.walkers or 65536, sweeps=args.sweeps or 1000, cpu_sample=cpu,
                    log=lambda r: print(json.dumps(r), flush=True))


def hetero_main(args):
    from mole_amd import workloads

    def log(r):
        per = r.pop("per_system")
        print(json.dumps(r), flush=True)
        print(json.dumps({"workload": "hetero-per-system", "systems": per}), flush=True)

    kw = {}
    if args.lanes_table:
        # a fixed lanes-per-walker table (JSON {electrons: lanes}): no tuning,
        # so a profiler sees the same launches every run
        kw = dict(tune=False, table={int(k): int(v) for k, v in json.loads(args.lanes_table).items()},
                  chunk=args.chunk or 20)
    workloads.hetero(sweeps=args.sweeps or 100, total_walkers=args.walkers or (1 << 20), log=log, **kw)


def learned_main(args):
    from mole_amd import workloads

    for Z in (3, 4, 10):
        workloads.learned(Z=Z, walkers=args.walkers or 65536, sweeps=args.sweeps or 400,
                          log=lambda r: print(json.dumps(r), flush=True))


def learned_record(log_runs=False):
    """The metric's second half for the line: learned vs box on the Ne atom
    (training included, all on the device)."""
    from mole_amd import workloads

    def log(r):
        if log_runs:
            print(json.dumps(r), file=sys.stderr, flush=True)

    r = workloads.learned(Z=10, log=log)

    def brief(x):
        return {k: x[k] for k in ("step", "energy", "error", "error_tau", "acceptance", "tau_int_mean",
                                  "moves_per_s", "ess_per_s", "kappa")}
    return dict(system="Ne atom (10 electrons; 1s/2s/2p STO basis, seed 1)", walkers=r["walkers"],
                sweeps=r["sweeps"], precision=r["precision"], training=r["training"],
                box=[brief(b) for b in r["box"]], learned=brief(r["learned"]),
                ess_ratio_vs_best_box=r["ess_ratio_vs_best_box"],
                kappa_ratio_vs_best_box=r["kappa_ratio_vs_best_box"],
                best_box_side_ess=r["best_box_side_ess"], best_box_side_tau=r["best_box_side_tau"],
                tau_ratio_vs_best_box=r["tau_ratio_vs_best_box"],
                energy_diff_sigma=r["energy_diff_sigma"],
                energy_diff_sigma_vs_best_ess_box=r["energy_diff_sigma_vs_best_ess_box"],
                reference=r.get("reference"),
                energy_diff_sigma_vs_reference=r.get("energy_diff_sigma_vs_reference"),
                tau_estimator="scripts/statfor.rs:30-54 on the device (mole_tau_int)")


def cpu_dmc_sample(walkers=4096, steps=200):
    """The oracle's DmcRunner (one thread, as the reference's diffuse loop
    runs) on a bounded sample of a DMC workload."""
    from oracle import sj_ffi as SJ

    def sample(sj, eref, tau, brancher):
        wv = SJ.SjWave(sj)
        d = SJ.Dmc(wv, walkers=walkers, seed=5, tau=tau, reference_energy=eref, brancher=brancher)
        t0 = time.perf_counter()
        d.diffuse(tau, steps, 10, 0)
        dt = time.perf_counter() - t0
        return dict(value=walkers * steps * sj.n_electrons / dt, unit="walker-moves/s", cores=1, kind="port",
                    sample=f"{walkers} walkers x {steps} DMC steps on the oracle (parity algorithm: SM moves, "
                           f"cached local energies), 1 thread, {dt:.2f} s")
    return sample


def dmc_main(args):
    from mole_amd import workloads

    cpu = None if args.no_cpu_baseline else cpu_dmc_sample()
    workloads.dmc(walkers=args.walkers or 65536, iters=args.sweeps or 1000, cpu_sample=cpu,
                  log=lambda r: print(json.dumps(r), flush=True))


def launch_main(args):
    """One fixed launch of the Slater-Jastrow sweep kernel, for profilers
    (kernel traces and counter passes need a short program whose last
    dispatch of the kernel is the measured one): the atom Z of the atoms
    workload, 4 warm-up sweeps, then one launch of --sweeps sampled sweeps.
    The learned proposal uses an untrained network (seed 2): its cost per move
    does not depend on the weights."""
    import mole_amd as M
    from mole_amd import systems

    prop = {"box": M.PROPOSAL_BOX, "diffuse": M.PROPOSAL_DIFFUSE, "learned": M.PROPOSAL_LEARNED}[args.proposal]
    prec = M.PREC_MIXED if args.precision == "mixed" else M.PREC_FP64
    sj = systems.atom(args.Z)
    W = args.walkers or 65536
    sweeps = args.sweeps or 20
    with M.Context([M.slater_jastrow(sj, W)], seed=1, step=args.side, block_size=2, precision=prec,
                   lu_period=64, proposal=prop) as ctx, M.Mlp(2) as net:
        if prop == M.PROPOSAL_LEARNED:
            net.attach(ctx)
        elif args.lanes:
            ctx.set_walker_lanes(args.lanes)
        ctx.init_walkers()
        ctx.sweep(4)
        ctx.synchronize()
        ctx.sweep(sweeps, sample=True)
        ctx.synchronize()
        ms, _ = ctx.last_timing()
        lanes = ctx.walker_lanes()
    print(json.dumps(dict(workload="launch", proposal=args.proposal, Z=args.Z, walkers=W, sweeps=sweeps, lanes=lanes,
                          precision=args.precision, ms=ms,
                          moves_per_s=W * sj.n_electrons * sweeps / (ms * 1e-3))), flush=True)


def copy_main(args):
    """The plain device copy (copy_probe_kernel) that calibrates the HBM byte
    counters: --bytes from one device buffer to another, --repeats times."""
    import mole_amd as M

    n = args.bytes // 8
    ms = M.debug_copy(n, args.repeats)
    print(json.dumps(dict(workload="copy", bytes_read=n * 8, bytes_written=n * 8, ms_per_copy=ms,
                          gb_per_s=2 * n * 8 / ms / 1e6)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["h2-sr", "atoms", "hetero", "learned", "dmc", "launch", "copy"],
                    default="h2-sr")
    ap.add_argument("--walkers", type=int, default=0)
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed SR iterations")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed SR iterations on a separate context (clocks settle)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="leave out the learned-vs-box record")
    ap.add_argument("--no-helium", action="store_true", help="leave out the helium singlet record")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--variant", choices=sorted(VARIANTS), default="auto")
    ap.add_argument("--host-loop", action="store_true",
                    help="one host round trip per SR iteration instead of the device-side loop")
    # --workload launch / copy (profiler runs)
    ap.add_argument("--proposal", choices=["box", "diffuse", "learned"], default="box")
    ap.add_argument("--Z", type=int, default=3)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--side", type=float, default=0.6)
    ap.add_argument("--lanes-table", default="", help="hetero: fixed lanes per walker, JSON {electrons: lanes}")
    ap.add_argument("--chunk", type=int, default=0, help="hetero with --lanes-table: sweeps per graph launch")
    ap.add_argument("--bytes", type=int, default=256 << 20)
    ap.add_argument("--repeats", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: walkers are independent Markov chains and this "
              "build measures one GPU's share; multi-GPU sharding is not built.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "atoms":
        atoms_main(args)
        return
    if args.workload == "hetero":
        hetero_main(args)
        return
    if args.workload == "learned":
        learned_main(args)
        return
    if args.workload == "dmc":
        dmc_main(args)
        return
    if args.workload == "launch":
        launch_main(args)
        return
    if args.workload == "copy":
        copy_main(args)
        return

    r = gpu_run(args.steps, args.warmup, variant=args.variant, host_loop=args.host_loop)
    moves = args.steps * WALKERS * SWEEPS * N_ELEC
    value = moves / r["seconds"]
    ms_step = 1e3 * r["seconds"] / args.steps
    # dominant kernel: one sweep launch of SWEEPS sweeps per iteration
    avg_kernel_s = 1e-3 * float(np.mean(r["kernel_ms"]))
    fl = flops_per_launch(WALKERS, SWEEPS, SWEEPS - BLOCK)
    achieved = fl / avg_kernel_s / 1e12
    # compute roofline: the kernel issues no MFMA; its FP64 work runs on the
    # vector ALUs (MI355X FP64 vector peak 78.6 TFLOP/s)
    roofline = dict(bound="fp64-valu", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP64_TFLOPS,
                    traffic=TRAFFIC_BYTES_PER_LAUNCH if args.variant in ("auto", "ws") else None,
                    traffic_source=TRAFFIC_SOURCE if args.variant in ("auto", "ws") else None,
                    kernel={"lane": "sweep2_lane_kernel<LCAO2>",
                            "quad": "sweep2_quad_kernel<LCAO2,2>"}.get(args.variant, "sweep2_ws_kernel<LCAO2,2,6>"),
                    kernel_ms=1e3 * avg_kernel_s,
                    flop_per_launch=fl)
    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpus()
        cpu = cpu_baseline(args.cpu_threads or host["threads"], host=host)
    learned = None if args.no_learned else learned_record()
    helium = None if args.no_helium else helium_record(args.cpu_threads, cpu=not args.no_cpu_baseline)
    e_last = r["energies"][-1]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Metropolis walkers from Philox seed 0)",
        "config": {
            "workload": "H2 LCAO sigma_g^2 x Pade Jastrow + stochastic reconfiguration",
            "walkers": WALKERS, "sweeps_per_iteration": SWEEPS, "block": BLOCK,
            "box_side": BOX, "sr_step": SR_STEP, "sr_shift": SR_EPS, "nuclei_z": [-0.7, 0.7],
            "parallelism": "single GPU, wave-specialised workgroups of 64 walkers",
            "optimizer_loop": "host" if args.host_loop else "device (mole_optimize)",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "result": {"energy": e_last[0], "error": e_last[1], "acceptance": e_last[2],
                   "params": r["params"], "energy_first": r["energies"][0][0]},
        "learned": learned,
        "helium": helium,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
