"""Autotuner for the Slater-Jastrow launch: lanes per walker per size class.

For every distinct electron count in a set of systems, one representative
system is timed alone on the device at each candidate L (lanes of a wave
sharing one walker) over a short run; the fastest L that fits the LDS budget
wins.  Results do not depend on L (every sum has a fixed order), so tuning
only moves time.  Tables can be saved as JSON and reused.
"""
import json
import os

import mole_amd as M

CANDIDATES = (1, 2, 4, 8, 16, 32, 64)


def time_lanes(system, walkers, lanes, *, sweeps=6, precision=M.PREC_MIXED, lu_period=64, seed=7,
               sample=True, lds_budget=None):
    """Device ms per sweep of `walkers` walkers of one system at L = lanes
    (None if the walker state does not fit the device, or lds_budget, at
    that L)."""
    dev = M.slater_jastrow(system, walkers)
    with M.Context([dev], seed=seed, step=0.6, block_size=2, precision=precision,
                   lu_period=lu_period) as ctx:
        try:
            ctx.set_walker_lanes(lanes)
        except M.MoleError as e:
            if e.status == M.ERR_UNSUPPORTED:
                return None
            raise
        if lds_budget is not None and ctx.launch_info()[1] > lds_budget:
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
    reps = {}
    for s in systems:
        reps.setdefault(s.n_electrons, s)
    for N, s in sorted(reps.items()):
        W = walkers_of(s)
        best, best_ms = None, None
        for L in candidates:
            ms = time_lanes(s, W, L, sweeps=sweeps, precision=precision, lu_period=lu_period,
                            lds_budget=lds_budget)
            timings[f"{N}:{L}"] = ms
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
    return table, timings


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
