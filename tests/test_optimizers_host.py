"""The stateful optimizers of the C ABI (mole_optimizer_init /
mole_optimizer_update, host code shared with the device loop) against the
oracle's restatement of MomentumDescent, NesterovMomentum and OnlineLbfgs
(src/optimize/src/optimizers.rs:32-179): the energy gradient from the sums
against util.rs:6-46's formula, and the updates bit for bit over a sequence
of calls."""
import ctypes as C

import numpy as np
import pytest

import mole_amd as M
from oracle import oracle_ffi as O


def _sums(rng, np_, n=5000):
    # sums of a synthetic sample set: E_k, O_k (d ln psi/dp)
    E = rng.normal(-1.1, 0.3, n)
    Ok = rng.normal(0.2, 0.5, (n, np_)) + 0.3 * (E[:, None] + 1.1)
    s = M.SrSums()
    s.n_params = np_
    s.n = n
    s.sum_e = float(E.sum())
    for i in range(np_):
        s.sum_o[i] = float(Ok[:, i].sum())
        s.sum_oe[i] = float((Ok[:, i] * E).sum())
        for j in range(np_):
            s.sum_oo[i * np_ + j] = float((Ok[:, i] * Ok[:, j]).sum())
    return s


def _grad(s):
    np_ = s.n_params
    ebar = s.sum_e / s.n
    so = np.array(s.sum_o[:np_])
    soe = np.array(s.sum_oe[:np_])
    return 2.0 * (soe - ebar * so) / s.n


CASES = [
    (M.OPT_MOMENTUM, O.OPT_MOMENTUM, dict(momentum=0.9), dict(mu=0.9)),
    (M.OPT_NESTEROV, O.OPT_NESTEROV, dict(momentum=0.7), dict(mu=0.7)),
    (M.OPT_LBFGS, O.OPT_LBFGS, dict(history=1), dict(history=1)),
    (M.OPT_LBFGS, O.OPT_LBFGS, dict(history=3), dict(history=3)),
    (M.OPT_LBFGS, O.OPT_LBFGS, dict(history=16), dict(history=16)),
]


@pytest.mark.parametrize("kind,okind,kw,okw", CASES)
@pytest.mark.parametrize("np_", [1, 3, 5])
def test_update_sequence_matches_oracle(kind, okind, kw, okw, np_):
    rng = np.random.default_rng(100 * kind + np_)
    step = 0.05
    opt = M.Optimizer(kind, np_, step, **kw)
    ref = O.OracleOptimizer(okind, np_, step, **okw)
    pars = rng.normal(size=np_)
    for call in range(12):
        s = _sums(rng, np_)
        dp, g = opt.compute_parameter_update(pars, s)
        assert np.allclose(g, _grad(s), rtol=1e-13, atol=1e-15)
        dp_ref = ref.update(pars, g)
        assert np.array_equal(dp.view(np.uint64), dp_ref.view(np.uint64)), (call, dp, dp_ref)
        pars = pars + dp
    assert opt.state.iter == 12


def test_sd_and_sr_through_the_stateful_entry_point():
    # the same kinds through mole_optimizer_update equal mole_compute_parameter_update
    rng = np.random.default_rng(3)
    s = _sums(rng, 3)
    for kind in (M.OPT_SD, M.OPT_SR):
        opt = M.Optimizer(kind, 3, 0.05, sr_mode=M.SR_TEXTBOOK, sr_eps=1e-3)
        dp, g = opt.compute_parameter_update(np.zeros(3), s)
        dp0, g0 = M.parameter_update(s, kind, 0.05, M.SR_TEXTBOOK, 1e-3)
        assert np.array_equal(dp, dp0) and np.array_equal(g, g0)


def test_validation():
    with pytest.raises(M.MoleError) as e:
        M.Optimizer(M.OPT_LBFGS, 3, 0.1, history=0)
    assert e.value.status == M.ERR_INVALID_ARGUMENT
    with pytest.raises(M.MoleError):
        M.Optimizer(M.OPT_LBFGS, 3, 0.1, history=M.LBFGS_MAX_HISTORY + 1)
    with pytest.raises(M.MoleError):
        M.Optimizer(7, 3, 0.1)
    with pytest.raises(M.MoleError) as e:
        M.Optimizer(M.OPT_MOMENTUM, M.MAX_PARAMS + 1, 0.1)
    assert e.value.status == M.ERR_SHAPE
    opt = M.Optimizer(M.OPT_MOMENTUM, 2, 0.1, momentum=0.5)
    with pytest.raises(M.MoleError) as e:
        opt.compute_parameter_update(np.zeros(3), _sums(np.random.default_rng(0), 3))
    assert e.value.status == M.ERR_SHAPE
    # a corrupted state is refused rather than indexed
    opt = M.Optimizer(M.OPT_LBFGS, 2, 0.1, history=2)
    opt.state.head = 5
    with pytest.raises(M.MoleError):
        opt.compute_parameter_update(np.zeros(2), _sums(np.random.default_rng(0), 2))
    # LBFGS needs the parameters
    opt = M.Optimizer(M.OPT_LBFGS, 2, 0.1, history=2)
    with pytest.raises(M.MoleError):
        M.check(M.lib().mole_optimizer_update(C.byref(opt.state), C.byref(_sums(np.random.default_rng(0), 2)), None,
                                              M._dp(np.zeros(2)), None))
