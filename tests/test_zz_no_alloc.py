"""The allocation-free tau entry point, mole_tau_int_noalloc, through the C
ABI against mole_tau_int: two calls in a row (the second on a smaller
system and a shorter window) give the allocating entry point's values bit
for bit, although the library's static scratch keeps the first call's
values; a system larger than the scratch is refused before any device
work."""
import numpy as np
import pytest

import mole_amd as M
from mole_amd import systems as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def taus(W, steps, n, first, seed):
    with M.Context([M.slater_jastrow(S.atom(3), W)], seed=seed, step=0.6, block_size=10, flags=M.FLAG_TRACE,
                   record_sweeps=steps, lu_period=16) as ctx:
        ctx.init_walkers()
        ctx.run(steps)
        a = ctx.tau_int(n, first=first)
        b = ctx.tau_int_noalloc(n, first=first)
    return a, b


def test_noalloc_tau_matches_allocating_entry_point():
    a1, b1 = taus(4096, 130, 120, 0, seed=1)
    a2, b2 = taus(700, 60, 30, 15, seed=2)  # smaller: the scratch still holds the first call's 4,096 values
    assert np.array_equal(bits(a1), bits(b1))
    assert np.array_equal(bits(a2), bits(b2))
    assert b2.size == 700 and np.all(b2 >= 1.0)


def test_noalloc_refuses_more_walkers_than_its_scratch():
    W = M.TAU_NOALLOC_MAX_WALKERS + 1
    with M.Context([M.helium(W)], seed=0, step=0.5, block_size=1, flags=M.FLAG_TRACE, record_sweeps=3) as ctx:
        ctx.init_walkers()
        ctx.run(3)  # two samples: the window is valid, the walker count is not
        with pytest.raises(M.MoleError) as ei:
            ctx.tau_int_noalloc(2)
        assert ei.value.status == M.ERR_UNSUPPORTED
