"""Statistics of walker series on the host: the integrated autocorrelation
time of scripts/statfor.rs:30-54 (scripts/statfor.py:22-40), vectorised over
walkers, and the effective-sample and efficiency figures built on it."""
import numpy as np

MAX_LAG = 200


def tau_int(x, max_lag=MAX_LAG):
    """tau = max(1, 1 + 2 sum_{i=1..min(max_lag, n-1)} rho_i), the sum cut at
    the first negative rho_i; rho_i = sum_t (x_t - m)(x_{t+i} - m) /
    ((n - i) var), var with ddof 1.  x: (n,) or (walkers, n); returns one tau
    per series."""
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    n = x.shape[1]
    d = x - x.mean(axis=1, keepdims=True)
    var = np.sum(d * d, axis=1) / (n - 1)
    L = min(max_lag, n - 1)
    # lagged sums of products by FFT (zero padded: no wrap-around)
    size = 1 << int(np.ceil(np.log2(2 * n)))
    f = np.fft.rfft(d, size, axis=1)
    acov = np.fft.irfft(f * np.conj(f), size, axis=1)[:, 1:L + 1]
    rho = acov / (var[:, None] * (n - np.arange(1, L + 1))[None, :])
    keep = np.cumprod(rho >= 0.0, axis=1)
    tau = 1.0 + 2.0 * np.sum(rho * keep, axis=1)
    return np.maximum(tau, 1.0)


def blocking_error(block_sum, block_sum2, n_blocks):
    """sqrt((<B^2> - <B>^2)/(n_b - 1)) of block means (src/vmc/src/vmc.rs:150-170)."""
    mb = block_sum / n_blocks
    return float(np.sqrt(max(block_sum2 / n_blocks - mb * mb, 0.0) / (n_blocks - 1)))
