"""One fixed launch of the Slater-Jastrow sweep kernel, for profilers
(rocprofv3 kernel traces and counter passes need a short program whose last
dispatch of the kernel is the measured one):

    python -m mole_amd.probe --proposal learned --Z 3 --sweeps 20

builds the atom Z of the atoms workload with `walkers` walkers, runs 4 warm-up
sweeps, then one launch of `sweeps` sampled sweeps, and prints its device
time.  The learned proposal uses an untrained network (seed 2): its cost per
move does not depend on the weights."""
import argparse
import json

import mole_amd as M
from mole_amd import systems


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--proposal", choices=["box", "diffuse", "learned"], default="box")
    ap.add_argument("--Z", type=int, default=3)
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--step", type=float, default=0.6)
    a = ap.parse_args(argv)
    prop = {"box": M.PROPOSAL_BOX, "diffuse": M.PROPOSAL_DIFFUSE, "learned": M.PROPOSAL_LEARNED}[a.proposal]
    prec = M.PREC_MIXED if a.precision == "mixed" else M.PREC_FP64
    sj = systems.atom(a.Z)
    with M.Context([M.slater_jastrow(sj, a.walkers)], seed=1, step=a.step, block_size=2, precision=prec,
                   lu_period=64, proposal=prop) as ctx, M.Mlp(2) as net:
        if prop == M.PROPOSAL_LEARNED:
            net.attach(ctx)
        else:
            ctx.set_walker_lanes(a.lanes)
        ctx.init_walkers()
        ctx.sweep(4)
        ctx.synchronize()
        ctx.sweep(a.sweeps, sample=True)
        ctx.synchronize()
        ms, _ = ctx.last_timing()
    print(json.dumps(dict(proposal=a.proposal, Z=a.Z, walkers=a.walkers, sweeps=a.sweeps, ms=ms,
                          moves_per_s=a.walkers * sj.n_electrons * a.sweeps / (ms * 1e-3))), flush=True)


if __name__ == "__main__":
    main()
