"""Shared fixtures of the learned-proposal tests (CPU oracle and GPU)."""
import numpy as np

from oracle import mlp_ffi as ML


def shaped_weights():
    """Xavier weights with a small shift and a strongly position-dependent
    width: an asymmetric proposal that still mixes fast."""
    w = ML.xavier(2)
    W4 = w[ML.OFF["W4"]:ML.OFF["W4"] + 256].reshape(64, 4)
    W4[:, :3] *= 0.1
    W4[:, 3] *= 2.0
    w[ML.OFF["W4"]:ML.OFF["W4"] + 256] = W4.ravel()
    return w


def hydrogen_radial_chi2(r, n_bins=30, r_max=4.0):
    """chi^2 of |x| against the exact 1s radial density 4 r^2 exp(-2 r)."""
    edges = np.linspace(0.0, r_max, n_bins + 1)
    cdf = lambda x: 1.0 - np.exp(-2 * x) * (1 + 2 * x + 2 * x * x)
    p = np.append(np.diff(cdf(edges)), 1.0 - cdf(r_max))
    h = np.append(np.histogram(r, edges)[0], np.sum(r >= r_max))
    e = p * r.size
    return np.sum((h - e) ** 2 / e), n_bins


def exact_hydrogen_configs(W, seed):
    rng = np.random.default_rng(seed)
    r = rng.gamma(3.0, 0.5, W)
    d = rng.normal(size=(W, 3))
    return d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
