"""The statfor text format and analysis (mole_amd/statfor.py) against
scripts/statfor.rs: the reader's first-column / last-line-dropped rule, the
correlation time against the oracle's restatement (oracle/vmc.c
orc_tau_corr, statfor.rs:30-54) and the vectorised estimator, variance with
ddof 1, and the blocking table against closed forms (iid: sigma/sqrt(n) at
every block size; AR(1): the plateau sqrt(var tau / n), tau =
(1+rho)/(1-rho))."""
import io
import os
import subprocess
import sys

import numpy as np

from mole_amd import statfor, stats
from oracle import oracle_ffi as O

ROOT = os.path.join(os.path.dirname(__file__), "..")


def ar1(n, rho, seed):
    rng = np.random.default_rng(seed)
    x = np.empty(n)
    x[0] = rng.normal()
    e = rng.normal(size=n) * np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


def test_format_round_trip_and_last_line_rule():
    x = np.random.default_rng(0).normal(size=37)
    f = io.StringIO()
    statfor.write_series(f, x)
    text = f.getvalue()
    assert len(text.splitlines()) == 38
    assert np.array_equal(statfor.read_series(io.StringIO(text)), x)
    # the reader takes the first whitespace field and drops the last line
    assert np.array_equal(statfor.read_series(io.StringIO("1.5 junk\n2.5\t7\n3.5\n")), [1.5, 2.5])


def test_analysis_against_oracle_and_numpy():
    x = ar1(5000, 0.8, 1)
    r = statfor.analyse(x)
    assert abs(r["average"] - x.mean()) < 1e-13
    assert abs(r["variance"] - np.var(x, ddof=1)) < 1e-12 * r["variance"]
    t_orc = O.lib().orc_tau_corr(O.dptr(np.ascontiguousarray(x)), x.size)
    assert abs(r["tcorr"] - t_orc) < 1e-10 * t_orc
    assert abs(r["tcorr"] - stats.tau_int(x)[0]) < 1e-9 * t_orc
    assert abs(r["n_eff"] - x.size / r["tcorr"]) < 1e-9
    assert abs(r["sigma"] - np.sqrt(r["variance"] * r["tcorr"] / x.size)) < 1e-15
    # AR(1): tau about (1+rho)/(1-rho) = 9 (the truncated estimator)
    assert 6.0 < r["tcorr"] < 12.0


def test_blocking_table_closed_forms():
    n = 40000
    x = np.random.default_rng(2).normal(size=n)
    rows = statfor.blocking_table(x)
    sizes = [s for s, _ in rows]
    assert sizes[0] == 1 and sizes[1] - sizes[0] == max(n // 20 // 100, 1)
    assert max(sizes) <= n // 20
    err1 = rows[0][1]
    assert abs(err1 - x.std(ddof=0) / np.sqrt(n - 1)) < 1e-12
    # iid: no growth with block size beyond noise.  statfor's chunks keep a
    # short last chunk whose noisy mean inflates a row (the reference's own
    # behaviour), so only rows whose last chunk is at least half a block
    # are held to the closed form
    good = [(s, e) for s, e in rows if n % s == 0 or n % s >= s // 2]
    assert len(good) >= 20
    for s, e in good:
        assert e < 1.6 / np.sqrt(n), s
    assert any(e > 1.6 / np.sqrt(n) for s, e in rows if 0 < n % s < s // 10)
    y = ar1(n, 0.9, 3)
    ry = statfor.blocking_table(y)
    plateau = np.sqrt(np.var(y, ddof=1) * 19.0 / n)
    big = [e for s, e in ry if s >= 200 and (n % s == 0 or n % s >= s // 2)]
    assert 0.6 * plateau < np.median(big) < 1.4 * plateau
    assert ry[0][1] < 0.5 * plateau  # block size 1 underestimates correlated errors


def test_cli_prints_statfor_lines(tmp_path):
    x = ar1(3000, 0.5, 4)
    buf = io.StringIO()
    statfor.write_series(buf, x)
    r = subprocess.run([sys.executable, "-m", "mole_amd.statfor"], input=buf.getvalue(), capture_output=True,
                       text=True, cwd=tmp_path, env=dict(os.environ, PYTHONPATH=ROOT), timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert [ln.split("\t")[0] for ln in lines] == ["average", "variance", "tcorr", "n eff", "sigma"]
    a = statfor.analyse(x)
    assert float(lines[2].split()[-1]) == a["tcorr"]
    assert (tmp_path / "corr.out").exists() and (tmp_path / "blocking.out").exists()
    assert len((tmp_path / "corr.out").read_text().splitlines()) == 200
