"""Autotuner for the Slater-Jastrow launch.

Stage 1 (tune): for every distinct electron count in a set of systems, one
representative system is timed alone on the device at each candidate L
(lanes of a wave sharing one walker: the wavefront packing, 64/L walkers per
wave) over a short run, and the LDS footprint of each L is recorded; the
fastest L that fits an LDS budget wins.  Stage 2 (tune_batch): the whole
batch, on its own context, is timed under the tables of several LDS budgets
(the budget caps the footprint the largest systems run at, hence the
occupancy of their launches), both tile schedules (one launch per footprint
class, or one persistent launch pulling longest-first tiles from a device
ticket), tables that take the fewest lanes within a tolerance of each
count's best single-system time, and several sweeps per graph-captured
launch; the fastest
combination is applied.  Results do not depend on any of these choices
(every sum has a fixed order), so tuning only moves time.  Tables can be
saved as JSON and reused.
"""
import json
import os
import time

import mole_amd as M

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)
BUDGETS = (16 << 10, 24 << 10, 32 << 10, 40 << 10, 64 << 10, None)
TOLERANCES = (0.05, 0.15)


def time_lanes(system, walkers, lanes, *, sweeps=6, precision=M.PREC_MIXED, lu_period=64, seed=7,
               sample=True, lds_budget=None, lds_out=None):
    """Device ms per sweep of `walkers` walkers of one system at L = lanes
    (None if the walker state does not fit the device, or lds_budget, at
    that L).  lds_out (a list) receives the workgroup LDS bytes at L."""
    dev = M.slater_jastrow(system, walkers)
    with M.Context([dev], seed=seed, step=0.6, block_size=2, precision=precision,
                   lu_period=lu_period) as ctx:
        try:
            ctx.set_walker_lanes(lanes)
        except M.MoleError as e:
            if e.status == M.ERR_UNSUPPORTED:
                return None
            raise
        lds = ctx.launch_info()[1]
        if lds_out is not None:
            lds_out.append(lds)
        if lds_budget is not None and lds > lds_budget:
            return None
        ctx.init_walkers()
        ctx.sweep(2, sample=sample)  # warm
        ctx.synchronize()
        ctx.sweep(sweeps, sample=sample)
        ms, _ = ctx.last_timing()
    return ms / sweeps


def tune(systems, walkers_of, *, candidates=CANDIDATES, sweeps=6, precision=M.PREC_MIXED,
         lu_period=64, lds_budget=None, verbose=False):
    """Best L per electron count.  walkers_of(system) gives the walker count
    to time with; lds_budget caps the LDS of a workgroup (the tiles of one
    4 KiB footprint class share a launch and its largest footprint, so the
    cap bounds the occupancy the largest systems run at)."""
    table = {}
    timings = {}
    footprint = {}
    reps = {}
    for s in systems:
        reps.setdefault(s.n_electrons, s)
    for N, s in sorted(reps.items()):
        W = walkers_of(s)
        best, best_ms = None, None
        for L in candidates:
            lds = []
            ms = time_lanes(s, W, L, sweeps=sweeps, precision=precision, lu_period=lu_period,
                            lds_budget=lds_budget, lds_out=lds)
            timings[f"{N}:{L}"] = ms
            if lds:
                footprint[f"{N}:{L}"] = lds[0]
            if ms is not None and (best_ms is None or ms < best_ms):
                best, best_ms = L, ms
        if best is None:  # nothing under the budget: the smallest footprint that fits the device
            fits = [c for c in candidates
                    if time_lanes(s, W, c, sweeps=1, precision=precision, lu_period=lu_period) is not None]
            if not fits:
                raise M.MoleError(M.ERR_UNSUPPORTED,
                                  f"autotune: no lanes per walker in {tuple(candidates)} fits the walker state of "
                                  f"the {N}-electron system in the device's LDS")
            best = max(fits)
        table[N] = best
        if verbose:
            timing = f"{best_ms:.3f} ms/sweep" if best_ms is not None else "over the LDS budget, smallest footprint"
            print(f"N={N:2d}: L={best} ({timing} for {W} walkers)", flush=True)
    timings["_lds"] = footprint
    return table, timings


def table_for_budget(timings, budget):
    """Best L per electron count among the timed ones whose LDS footprint is
    within budget (None: no cap); counts with nothing under the budget take
    their smallest-footprint L."""
    lds = timings.get("_lds", {})
    runs = {}
    for key, ms in timings.items():
        if key.startswith("_") or ms is None:
            continue
        N, L = (int(v) for v in key.split(":"))
        runs.setdefault(N, []).append((L, ms, lds.get(key, 0)))
    out = {}
    for N, rs in runs.items():
        fit = [r for r in rs if budget is None or r[2] <= budget]
        out[N] = min(fit, key=lambda r: r[1])[0] if fit else min(rs, key=lambda r: r[2])[0]
    return out


def table_within(timings, tol):
    """Per electron count, the fewest lanes whose single-system time is
    within (1 + tol) of that count's best.  Fewer lanes put more walkers in a
    wave; in a full batch, where every CU already has work, that can beat the
    lane count that wins on one system alone."""
    runs = {}
    for key, ms in timings.items():
        if key.startswith("_") or ms is None:
            continue
        N, L = (int(v) for v in key.split(":"))
        runs.setdefault(N, []).append((L, ms))
    out = {}
    for N, rs in runs.items():
        best = min(ms for _, ms in rs)
        out[N] = min(L for L, ms in rs if ms <= (1.0 + tol) * best)
    return out


def _time_config(ctx, systems, table, schedule, chunk, reps):
    apply(ctx, systems, table)
    ctx.set_tile_schedule(schedule)
    g = ctx.capture(chunk, sample=True)
    try:
        g.launch(1)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        g.launch(reps)
        ctx.synchronize()
        dt = time.perf_counter() - t0
    finally:
        g.close()
    return dt / (reps * chunk)


def tune_batch(ctx, systems, timings, *, budgets=BUDGETS, schedules=(M.SCHEDULE_CLASSES, M.SCHEDULE_PERSISTENT),
               tolerances=TOLERANCES, chunks=(5, 10, 20), reps=2, base_chunk=5, log=None):
    """Whole-batch tuning on an initialised context: seconds per sweep under
    every (LDS budget, schedule) pair at base_chunk sweeps per launch, then
    the fewer-lanes tables of table_within(tolerance) under the best
    schedule, then the best configuration at each chunk.  Applies the winner
    (lanes and schedule) to ctx and returns (best, trials) with best =
    dict(budget, schedule, chunk, table, s_per_sweep) (budget None and a
    tolerance key for a fewer-lanes table)."""
    trials = []
    tables = {}
    for b in budgets:
        t = table_for_budget(timings, b)
        key = tuple(sorted(t.items()))
        if key in tables.values():  # same table as a smaller budget
            continue
        tables[b] = key
    best = None
    for b, key in tables.items():
        table = dict(key)
        for sch in schedules:
            try:
                sps = _time_config(ctx, systems, table, sch, base_chunk, reps)
            except M.MoleError as e:  # e.g. the persistent schedule is not eligible
                trials.append(dict(budget=b, schedule=int(sch), chunk=base_chunk, error=str(e)))
                continue
            rec = dict(budget=b, schedule=int(sch), chunk=base_chunk, s_per_sweep=sps)
            trials.append(rec)
            if log:
                log(rec)
            if best is None or sps < best["s_per_sweep"]:
                best = dict(rec, table=table)
    for tol in tolerances:
        table = table_within(timings, tol)
        key = tuple(sorted(table.items()))
        if key in tables.values():
            continue
        tables[f"tol{tol}"] = key
        sps = _time_config(ctx, systems, table, best["schedule"], base_chunk, reps)
        rec = dict(budget=None, tolerance=tol, schedule=best["schedule"], chunk=base_chunk, s_per_sweep=sps)
        trials.append(rec)
        if log:
            log(rec)
        if sps < best["s_per_sweep"]:
            best = dict(rec, table=table)
    for c in chunks:
        if c == base_chunk:
            continue
        sps = _time_config(ctx, systems, best["table"], best["schedule"], c, max(1, reps * base_chunk // c))
        rec = {k: best[k] for k in ("budget", "tolerance") if k in best}
        rec.update(schedule=best["schedule"], chunk=c, s_per_sweep=sps)
        trials.append(rec)
        if log:
            log(rec)
        if sps < best["s_per_sweep"]:
            best = dict(rec, table=best["table"])
    apply(ctx, systems, best["table"])
    ctx.set_tile_schedule(best["schedule"])
    return best, trials


def apply(ctx, systems, table):
    for k, s in enumerate(systems):
        L = table.get(s.n_electrons) or table.get(str(s.n_electrons))
        if L:
            ctx.set_walker_lanes(int(L), sys=k)


def save(table, path):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1, sort_keys=True)


def load(path):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}
