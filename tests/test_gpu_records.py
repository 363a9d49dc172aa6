"""Per-block energy records streamed from the device (mole_read_blocks):
the ensemble form of what Runner::run logs once per block
(src/montecarlo/src/montecarlo.rs:31-42) and of VmcRunner's block means
(src/vmc/src/vmc.rs:150-170).  Checked against numpy reductions of the
local-energy trace, against the accumulators' block sums, against the
oracle's local energies, and for incremental reads during a run."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import statfor
from mole_amd import systems as S
from oracle import sj_ffi as SJ

pytestmark = pytest.mark.gpu


def blocks_from_trace(el, B):
    W, n = el.shape
    nb = n // B
    bm = np.array([[sum(el[w, b * B:(b + 1) * B]) * (1.0 / B) for b in range(nb)] for w in range(W)])
    return bm.sum(axis=0), (bm * bm).sum(axis=0)


def test_blocks_slater_jastrow_against_trace_stats_and_oracle():
    W, steps, B = 64, 60, 10
    sj = S.atom(3)
    with M.Context([M.slater_jastrow(sj, W)], seed=0, step=0.6, block_size=B, flags=M.FLAG_TRACE,
                   record_sweeps=steps, lu_period=16) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        rec = ctx.read_blocks()
        _, el = ctx.trace(steps, steps - B)
        st = ctx.stats()
    assert rec.size == (steps - B) // B
    assert np.array_equal(rec["block"], np.arange(rec.size)) and np.all(rec["n_walkers"] == W)
    s1, s2 = blocks_from_trace(el, B)
    assert np.allclose(rec["sum_block_mean"], s1, rtol=1e-13)
    assert np.allclose(rec["sum_block_mean2"], s2, rtol=1e-13)
    assert np.allclose(rec["mean"], s1 / W, rtol=1e-13)
    # the accumulators hold the same block means, summed in another order
    assert abs(rec["sum_block_mean"].sum() - st["sum_block_mean"]) < 1e-11 * abs(st["sum_block_mean"])
    assert abs(rec["sum_block_mean2"].sum() - st["sum_block_mean2"]) < 1e-11 * st["sum_block_mean2"]
    # the oracle's local energies of the same run
    o = SJ.run(SJ.SjWave(sj), seed=0, walkers=W, steps=steps, block_size=B, box_side=0.6, lu_period=16,
               n_threads=8)
    o1, _ = blocks_from_trace(o["e_local"], B)
    assert np.allclose(rec["sum_block_mean"], o1, rtol=1e-10)
    # per-walker dumps in the statfor format survive the round trip
    import io
    f = io.StringIO()
    statfor.write_series(f, el[0])
    assert np.array_equal(statfor.read_series(io.StringIO(f.getvalue())), el[0])


def test_blocks_two_electron_and_incremental_reads():
    W, B = 256, 10
    with M.Context([M.h2(W, 1.0, 0.5, 0.3)], seed=4, step=0.5, block_size=B, flags=M.FLAG_TRACE,
                   record_sweeps=80) as ctx:
        ctx.init_walkers()
        ctx.reset_stats()
        seen = []
        for k in range(8):  # one launch per block, reading the new record after each
            ctx.sweep(B, sample=True)
            new = ctx.read_blocks(first=len(seen))
            assert new.size == 1 and new["block"][0] == k
            seen.append(new[0])
        assert ctx.read_blocks(first=8).size == 0
        full = ctx.read_blocks()
        _, el = ctx.trace(80, 80)
    assert np.array_equal(np.array([r["sum_block_mean"] for r in seen]), full["sum_block_mean"])
    s1, s2 = blocks_from_trace(el, B)
    assert np.allclose(full["sum_block_mean"], s1, rtol=1e-13)
    assert np.allclose(full["sum_block_mean2"], s2, rtol=1e-13)


def test_blocks_need_a_trace():
    with M.Context([M.h2(16, 1.0, 0.5, 0.3)], seed=4, step=0.5, block_size=5) as ctx:
        ctx.init_walkers()
        ctx.run(20)
        with pytest.raises(M.MoleError) as ei:
            ctx.read_blocks()
        assert ei.value.status == M.ERR_DATA_ACCESS


# ---- integrated autocorrelation times on the device (scripts/statfor.rs:30-54)

def test_tau_int_on_device_matches_oracle_fold():
    # every walker's tau from the trace on the device, against the oracle's
    # sequential fold over the same series (bit for bit) and the numpy FFT
    # estimator; a window of the series as well; short series (n - 1 < 200
    # lags) and the first-negative cut are both exercised
    from mole_amd.stats import tau_int
    from oracle import oracle_ffi as O

    W, steps, B = 96, 260, 10
    sj = S.atom(4)
    with M.Context([M.slater_jastrow(sj, W)], seed=3, step=0.6, block_size=B, flags=M.FLAG_TRACE,
                   record_sweeps=steps, lu_period=16) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        n = steps - B
        _, el = ctx.trace(n, n)
        tau = ctx.tau_int(n)
        tau_w = ctx.tau_int(40, first=100)
        with pytest.raises(M.MoleError):
            ctx.tau_int(n + 1)
        with pytest.raises(M.MoleError):
            ctx.tau_int(1)
    L = O.lib()
    want = np.array([L.orc_tau_corr(O.dptr(np.ascontiguousarray(el[w])), n) for w in range(W)])
    want_w = np.array([L.orc_tau_corr(O.dptr(np.ascontiguousarray(el[w, 100:140])), 40) for w in range(W)])
    assert np.array_equal(bits(tau), bits(want))
    assert np.array_equal(bits(tau_w), bits(want_w))
    assert np.allclose(tau, tau_int(el), rtol=1e-9)
    assert np.all(tau >= 1.0) and np.any(tau > 1.5)


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)
