"""The text format and analysis of scripts/statfor.rs, for per-walker series
dumped from the device (local energies from mole_read_trace, DMC step
energies from mole_dmc_read_steps).

Format (statfor.rs:6-14, 86): one value per line, the first whitespace
field of each line parsed as a float; the last line is dropped by the
reader, so write_series appends one extra line (a copy of the last value).

Analysis (statfor.rs:17-81): mean, variance with ddof 1, the integrated
correlation time (lags up to 200, terms zeroed from the first negative
autocorrelation on, at least 1), n_eff = n / tau, sigma = sqrt(var tau / n),
and the blocking table (block sizes 1..n/20 in steps of max(1, n/20/100),
error sqrt((<B^2> - <B>^2)/(n_blocks - 1)) over data.chunks(size), the last
chunk possibly short).

    python -m mole_amd.statfor < series.txt

prints the five lines statfor prints and writes corr.out and blocking.out in
the working directory.
"""
import sys

import numpy as np

from mole_amd import stats

MAX_STEPS = 200
MIN_LEFT = 20
NSIZES = 100


def write_series(f, x):
    """Write x in the statfor input format (text file object or path)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    if isinstance(f, str):
        with open(f, "w") as fh:
            return write_series(fh, x)
    for v in x:
        f.write(f"{float(v)!r}\n")
    if x.size:
        f.write(f"{float(x[-1])!r}\n")  # dropped by the reader


def read_series(f):
    """read_data and the slice data[..len-1] of statfor.rs:6-14, 86."""
    if isinstance(f, str):
        with open(f) as fh:
            return read_series(fh)
    vals = [float(line.split()[0]) for line in f if line.strip()]
    return np.asarray(vals[:-1], dtype=np.float64)


def correlation_table(x, mean, variance):
    """(lag, rho) rows and tau of correlation() (statfor.rs:30-54)."""
    n = x.size
    max_i = min(MAX_STEPS, n - 1)
    d = x - mean
    rows = []
    tau, f = 1.0, 1.0
    for i in range(1, max_i + 1):
        c = float(np.dot(d[: n - i], d[i:])) / (variance * (n - i))
        if c < 0.0:
            f = 0.0
        tau += 2.0 * c * f
        rows.append((i, c))
    return rows, max(tau, 1.0)


def blocking_table(x):
    """(size, error) rows of blocking() (statfor.rs:58-81)."""
    n = x.size
    large = n // MIN_LEFT
    step = max(large // NSIZES, 1)
    rows = []
    for size in range(1, large + 1, step):
        nblocks = n // size
        blocks = [x[k:k + size].mean() for k in range(0, n, size)]  # chunks: the last may be short
        ave_blks = np.asarray(blocks)
        ave_sq = float(np.mean(ave_blks ** 2))
        ave = float(np.mean(ave_blks))
        rows.append((size, float(np.sqrt((ave_sq - ave * ave) / (nblocks - 1)))))
    return rows


def analyse(x):
    """The five numbers statfor prints, plus its two tables."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    mean = float(x.sum() / n)
    var = float(np.sum((mean - x) ** 2) / (n - 1))
    corr, tau = correlation_table(x, mean, var)
    return dict(average=mean, variance=var, tcorr=tau, n_eff=n / tau, sigma=float(np.sqrt(var * tau / n)),
                correlation=corr, blocking=blocking_table(x))


def tau_fast(x):
    """tcorr for many series at once (FFT; the same estimator)."""
    return stats.tau_int(x, MAX_STEPS)


def main(argv=None):
    x = read_series(sys.stdin)
    r = analyse(x)
    with open("corr.out", "w") as f:
        for i, c in r["correlation"]:
            f.write(f"{i} {c!r}\n")
    with open("blocking.out", "w") as f:
        for size, e in r["blocking"]:
            f.write(f"{size} {e!r}\n")
    print(f"average\t\t{r['average']!r}")
    print(f"variance\t{r['variance']!r}")
    print(f"tcorr\t\t{r['tcorr']!r}")
    print(f"n eff\t\t{r['n_eff']!r}")
    print(f"sigma\t\t{r['sigma']!r}")


if __name__ == "__main__":
    main()
