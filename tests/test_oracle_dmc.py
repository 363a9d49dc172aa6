"""Pins of the oracle's diffusion Monte Carlo (oracle/sj.c: orc_branch_sr,
orc_branch_simple, orc_dmc_init, orc_dmc_diffuse), which restates
DmcRunner::diffuse (src/dmc/src/dmc.rs:75-202), SRBrancher and
SimpleBranching (src/dmc/src/branching.rs:7-91).

* The branchers against plain-Python restatements built differently
  (numpy searchsorted for WeightedChoice; Python list pop/append for the
  kill/birth/clone/remove sequence) on the same Philox draws, and against
  their distributions.
* Zero-variance cases: the hydrogen 1s function (E_L = -1/2) and the He
  product function at zeta = 2 under the ionic Hamiltonian (E_L = -4) give
  those energies at every step with weights that never change.
* Physics: He with a Pade-Jastrow trial function reaches the exact
  ground-state energy -2.903724 (fixed-node is exact for the nodeless He
  ground state) within the statistical error and a time-step allowance.
* update_energies (running mean, variance, E_ref mixing) against a
  restatement over the recorded step energies.
"""
import numpy as np
import pytest

from mole_amd import systems
from oracle import oracle_ffi as OF
from oracle import sj_ffi as S

HE_EXACT = -2.903724


def _draw(seed, sys, step, i):
    return int(OF.lib().orc_stream_u64(seed, sys, step & 0xFFFFFFFF, S.DMC_BRANCH_EPOCH, i))


def _below(d, n):
    return (d * n) >> 64


def _as_u32(x):
    if not (x > 0):
        return 0
    return 0xFFFFFFFF if x >= 4294967295.0 else int(x)


@pytest.mark.parametrize("W", [1, 7, 64, 300])
def test_branch_sr_matches_weighted_choice(W):
    rng = np.random.default_rng(W)
    seed, sys, step = 11, 2, 5
    w = rng.exponential(1.0, W)
    w[rng.random(W) < 0.2] = 0.0
    w[0] = max(w[0], 0.1)
    src, wo = S.branch(w, seed, sys, step, S.BRANCH_SR)
    norm = W / w.max()
    iw = np.array([_as_u32(x * norm) for x in w], dtype=np.uint64)
    cum = np.cumsum(iw)
    total = int(cum[-1])
    u = np.array([_below(_draw(seed, sys, step, j), total) for j in range(W)], dtype=np.uint64)
    expect = np.searchsorted(cum, u, side="right")
    assert np.array_equal(src, expect)
    assert np.all(iw[src] > 0)  # a walker of integer weight 0 is never chosen
    assert np.allclose(wo, w.sum() / W, rtol=1e-15)


def test_branch_sr_distribution():
    # chosen indices are distributed as integer weight / total
    W = 8
    w = np.array([1.0, 2.0, 0.5, 4.0, 0.0, 3.0, 1.5, 2.5])
    iw = np.floor(w * W / w.max())
    p = iw / iw.sum()
    counts = np.zeros(W)
    for step in range(4000):
        src, _ = S.branch(w, 3, 0, step, S.BRANCH_SR)
        counts += np.bincount(src, minlength=W)
    n = counts.sum()
    mask = p > 0
    chi2 = np.sum((counts[mask] - n * p[mask]) ** 2 / (n * p[mask]))
    assert counts[~mask].sum() == 0
    assert chi2 < 30.0  # 6 degrees of freedom


def _simple_restated(w, seed, sys, step):
    W = w.size
    walkers = list(range(W))
    to_kill, to_birth = [], []
    for i in range(W):
        u = OF.lib().orc_u64_standard(_draw(seed, sys, step, i))
        x = w[i] + u
        n = min(int(x) if x > 0 else 0, 3)
        if n == 0:
            to_kill.append(i)
        else:
            to_birth.extend([i] * (n - 1))
    new = list(walkers)
    for idx in reversed(to_kill):
        new.pop(idx)
    new.extend(to_birth)
    k = 0
    if W > len(new):
        excess = W - len(new)
        for k in range(excess):
            new.append(walkers[_below(_draw(seed, sys, step, W + k), W)])
    else:
        excess = len(new) - W
        for k in range(excess):
            new.pop(_below(_draw(seed, sys, step, W + k), len(new)))
    return np.array(new, dtype=np.uint32)


@pytest.mark.parametrize("W,scale", [(1, 1.0), (10, 1.0), (50, 3.0), (200, 0.6), (200, 2.2)])
def test_branch_simple_matches_list_restatement(W, scale):
    rng = np.random.default_rng(W + int(10 * scale))
    w = rng.uniform(0, scale, W)
    for step in range(5):
        src, wo = S.branch(w, 7, 1, step, S.BRANCH_SIMPLE)
        expect = _simple_restated(w, 7, 1, step)
        assert np.array_equal(src, expect), step
        assert np.array_equal(wo, w[src])


def test_branch_simple_identity_for_unit_weights():
    w = np.ones(100)
    src, wo = S.branch(w, 1, 0, 0, S.BRANCH_SIMPLE)
    assert np.array_equal(src, np.arange(100)) and np.all(wo == 1.0)
    # all weights zero: every walker dies and W uniform clones replace them
    src, wo = S.branch(np.zeros(100), 1, 0, 0, S.BRANCH_SIMPLE)
    assert src.size == 100 and np.all(wo == 0.0)


def test_branch_sr_refuses_zero_total():
    with pytest.raises(ValueError):
        S.branch(np.zeros(4), 1, 0, 0, S.BRANCH_SR)


@pytest.mark.parametrize("brancher", [S.BRANCH_SR, S.BRANCH_SIMPLE])
def test_zero_variance_hydrogen(brancher):
    wv = S.SjWave(systems.hydrogenic(1))
    d = S.Dmc(wv, walkers=40, seed=4, tau=0.05, reference_energy=-0.5, brancher=brancher)
    e, err, se = d.diffuse(0.05, 60, 10, 2)
    assert np.all(np.abs(se + 0.5) < 1e-12)
    assert np.all(np.abs(e + 0.5) < 1e-12) and np.all(err < 1e-12)
    assert np.all(d.weight == 1.0)
    assert abs(d.reference_energy + 0.5) < 1e-12


def test_zero_variance_helium_ionic():
    wv = S.SjWave(systems.helium_product(2.0, 0.0), hamiltonian=1)
    d = S.Dmc(wv, walkers=40, seed=5, tau=0.02, reference_energy=-4.0)
    e, err, se = d.diffuse(0.02, 40, 10, 1)
    assert np.all(np.abs(se + 4.0) < 1e-11)
    assert np.all(np.abs(d.weight - 1.0) < 1e-12)


def _update_energies_restated(step_energy, block_size, num_eq_blocks, eref):
    energies, vars_ = [], []
    blocks = step_energy.size // block_size
    for b in range(blocks):
        blk = step_energy[b * block_size:(b + 1) * block_size]
        if b == num_eq_blocks:
            energy = sum(blk) / block_size
            eref = (eref + energy) / 2.0
            energies.append(energy)
            vars_.append(0.0)
        if b > num_eq_blocks:
            energy = sum(blk) / block_size
            prev = energies[-1]
            k = b - num_eq_blocks
            energies.append(prev + (energy - prev) / k)
            eref = (eref + energies[-1]) / 2.0
            vars_.append(vars_[-1] + ((energy - prev) * (energy - energies[-1]) - vars_[-1]) / k)
    errs = [np.sqrt(v / (i + 1)) for i, v in enumerate(vars_)]
    return np.array(energies), np.array(errs), eref


@pytest.mark.parametrize("brancher", [S.BRANCH_SR, S.BRANCH_SIMPLE])
def test_helium_ground_state_and_block_bookkeeping(brancher):
    wv = S.SjWave(systems.helium_product(1.6875, 0.5, 0.3))
    eref0 = -2.88
    d = S.Dmc(wv, walkers=256, seed=3, tau=0.02, reference_energy=eref0, brancher=brancher)
    iters, block, eq = 3000, 100, 5
    e, err, se = d.diffuse(0.02, iters, block, eq)
    # the reference's bookkeeping (E_ref is only mixed from block eq on)
    e2, err2, eref = _update_energies_restated(se, block, eq, eref0)
    assert np.allclose(e, e2, rtol=1e-13) and np.allclose(err, err2, rtol=1e-10)
    assert abs(d.reference_energy - eref) < 1e-12
    assert e.size == iters // block - eq
    # the ground state: 3 sigma plus a time-step allowance of 0.01
    assert abs(e[-1] - HE_EXACT) < 3 * err[-1] + 0.01, (e[-1], err[-1])
    # the DMC energy is below the trial function's variational energy (-2.878)
    assert e[-1] < -2.885


def test_device_order_sum():
    # orc_device_sum follows the device's fixed order; it is a sum: within a
    # few ulps of the exact (fsum) value, and for n <= 256 it is one pairwise
    # tree over the inputs (each thread holds one element)
    import ctypes as C
    import math
    from oracle import oracle_ffi as OF
    L = OF.lib()
    L.orc_device_sum.restype = C.c_double
    L.orc_device_sum.argtypes = [C.POINTER(C.c_double), C.c_size_t]
    rng = np.random.default_rng(1)
    for n in (1, 7, 256, 257, 4096, 4097, 70000):
        x = np.ascontiguousarray(rng.exponential(1.0, n))
        s = L.orc_device_sum(OF.dptr(x), n)
        assert abs(s - math.fsum(x)) <= 4 * n * np.finfo(float).eps * math.fsum(x)
    x = np.ascontiguousarray(rng.normal(size=200))
    sh = np.zeros(256)
    sh[:200] = x
    off = 128
    while off:
        sh[:off] = sh[:off] + sh[off:2 * off]
        off //= 2
    assert L.orc_device_sum(OF.dptr(x), 200) == sh[0]
