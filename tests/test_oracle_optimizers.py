"""Pins of the oracle's stateful optimizers (oracle/vmc.c, orc_stateful_opt_*),
restating MomentumDescent, NesterovMomentum and OnlineLbfgs of
src/optimize/src/optimizers.rs:32-179.

Momentum and Nesterov are pinned by their closed forms on a constant
gradient.  The LBFGS two-loop recursion is pinned against the explicit BFGS
inverse-Hessian recursion H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T
over the same pair list, starting from H0 = gamma I, which is what the
two-loop recursion computes (Nocedal & Wright, Numerical Optimization,
algorithm 7.4): a swapped s/y, a pair taken in the wrong order or a wrong
gamma changes the result by far more than rounding.  The pair list itself
(each call pushes its pair twice, optimizers.rs:163,170; oldest dropped
beyond `history`) is rebuilt independently here.
"""
import numpy as np
import pytest

from oracle import oracle_ffi as O


def test_momentum_closed_form():
    g = np.array([0.3, -1.2, 2.5])
    eta, mu = 0.05, 0.8
    opt = O.OracleOptimizer(O.OPT_MOMENTUM, 3, eta, mu)
    for k in range(1, 8):
        dp = opt.update(np.zeros(3), g)
        # m_k = -k eta g, dp = mu m_k
        assert np.allclose(dp, mu * (-k * eta * g), rtol=1e-14, atol=0)


def test_nesterov_closed_form():
    g = np.array([0.7, -0.4])
    eta, mu = 0.1, 0.6
    opt = O.OracleOptimizer(O.OPT_NESTEROV, 2, eta, mu)

    def m(k):  # m_k = mu m_{k-1} + eta g, m_0 = 0
        return eta * g * (1 - mu**k) / (1 - mu)

    for k in range(1, 8):
        dp = opt.update(np.zeros(2), g)
        assert np.allclose(dp, -(mu * m(k - 1) + (1 + mu) * m(k)), rtol=1e-14, atol=1e-300)


def _bfgs_direction(pairs, gamma, g):
    n = g.size
    H = gamma * np.eye(n)
    for s, y in pairs:  # oldest first; the newest pair is applied last
        rho = 1.0 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return -H @ g


def _lbfgs_model(pars_seq, grad_seq, history):
    """Pair lists and scale factors of OnlineLbfgs, call by call."""
    pairs, out = [], []
    pp = np.full(pars_seq[0].size, 1e-5)
    gp = np.full(pars_seq[0].size, 1e-5)
    for it, (p, g) in enumerate(zip(pars_seq, grad_seq)):
        def push():  # update_curvature_pairs
            if len(pairs) >= history:
                pairs.pop(0)
            pairs.append((p - pp, g - gp))
        push()
        held = list(pairs)  # initial_direction runs between the two pushes
        if it == 0:
            gamma = 1e-10
        else:
            gamma = sum((s @ y) / (y @ y) for s, y in held) / min(it, history)
        out.append((held, gamma))
        push()
        pp, gp = p.copy(), g.copy()
    return out


@pytest.mark.parametrize("history", [1, 2, 3, 5, 8])
def test_lbfgs_matches_explicit_bfgs_recursion(history):
    rng = np.random.default_rng(history)
    n, step, calls = 4, 0.3, 9
    # a quadratic model: gradients of 0.5 p^T A p - b^T p at a parameter walk
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T + n * np.eye(n)
    b = rng.normal(size=n)
    opt = O.OracleOptimizer(O.OPT_LBFGS, n, step, history=history)
    pars = rng.normal(size=n)
    pars_seq, grad_seq, dps = [], [], []
    for _ in range(calls):
        g = A @ pars - b
        dp = opt.update(pars, g)
        pars_seq.append(pars.copy())
        grad_seq.append(g)
        dps.append(dp)
        pars = pars + dp
    model = _lbfgs_model(pars_seq, grad_seq, history)
    for it in range(calls):
        held, gamma = model[it]
        expect = -step * _bfgs_direction(held, gamma, grad_seq[it])
        assert np.allclose(dps[it], expect, rtol=1e-8, atol=1e-14 * np.abs(expect).max()), it


def test_explicit_recursion_satisfies_secant_condition():
    # checks the checker: the explicit form used above is BFGS, i.e. the
    # updated inverse Hessian maps the newest pair's y onto its s
    rng = np.random.default_rng(5)
    n = 3
    pairs = [(rng.normal(size=n), rng.normal(size=n)) for _ in range(3)]
    pairs = [(s, s + 0.1 * y) for s, y in pairs]  # s.y > 0
    s, y = pairs[-1]
    # _bfgs_direction(pairs, gamma, g) = -H g
    assert np.allclose(-_bfgs_direction(pairs, 0.7, y), s, rtol=1e-12)
    assert not np.allclose(-_bfgs_direction(pairs[::-1], 0.7, y), s, rtol=1e-3)


def test_planted_mistake_is_caught():
    # a direction built with s and y swapped (a plausible transposition)
    # is far from the oracle's, so the explicit-recursion test discriminates
    rng = np.random.default_rng(9)
    n = 4
    Q = rng.normal(size=(n, n))
    A = Q @ Q.T + n * np.eye(n)
    opt = O.OracleOptimizer(O.OPT_LBFGS, n, 0.3, history=3)
    pars = rng.normal(size=n)
    pars_seq, grad_seq, dps = [], [], []
    for _ in range(4):
        g = A @ pars
        dp = opt.update(pars, g)
        pars_seq.append(pars.copy())
        grad_seq.append(g)
        dps.append(dp)
        pars = pars + dp
    held, gamma = _lbfgs_model(pars_seq, grad_seq, 3)[3]
    swapped = [(y, s) for s, y in held]
    wrong = -0.3 * _bfgs_direction(swapped, gamma, grad_seq[3])
    assert not np.allclose(dps[3], wrong, rtol=1e-3)
