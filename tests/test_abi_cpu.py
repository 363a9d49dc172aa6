"""CPU checks of the C ABI: the library loads, exports every entry point the
headers declare, and rejects invalid descriptors before touching a device."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import mole_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mole_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(M.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    for n in sorted(names):
        assert hasattr(L, n), n
    # and the Python plumbing binds exactly the declared set
    assert set(M.exported_symbols()) == names


def test_status_strings():
    L = M.lib()
    assert L.mole_status_string(M.OK) == b"ok"
    assert L.mole_status_string(M.ERR_LINALG) == b"linear algebra error"
    assert L.mole_status_string(M.ERR_DATA_ACCESS) == b"data access error"
    assert L.mole_status_string(12345) == b"unknown status"


def _create(systems, **kw):
    with pytest.raises(M.MoleError) as ei:
        M.Context(systems, **kw)
    return ei.value.status


def test_validation_before_device():
    assert _create([M.helium(16)], step=0.0) == M.ERR_INVALID_ARGUMENT
    assert _create([M.helium(16)], step=float("nan")) == M.ERR_INVALID_ARGUMENT
    assert _create([M.helium(16)], block_size=0) == M.ERR_INVALID_ARGUMENT
    assert _create([M.helium(0)]) == M.ERR_INVALID_ARGUMENT
    assert _create([M.helium(16)], proposal=7) == M.ERR_UNSUPPORTED
    assert _create([M.helium(16)], o_convention=9) == M.ERR_INVALID_ARGUMENT
    assert _create([M.helium(16)], flags=M.FLAG_TRACE, record_sweeps=0) == M.ERR_INVALID_ARGUMENT
    bad = M.helium(16)
    bad.desc.n_params = 2
    assert _create([bad]) == M.ERR_SHAPE
    bad = M.helium(16)
    bad.desc.n_up = 2
    assert _create([bad]) == M.ERR_SHAPE
    bad = M.helium(16)
    bad.desc.wf_kind = 99
    assert _create([bad]) == M.ERR_UNSUPPORTED
    assert "kind 99" in M.lib().mole_last_error().decode()


def test_null_context_and_arguments():
    L = M.lib()
    assert L.mole_sweep(None, 1, 0) == M.ERR_INVALID_ARGUMENT
    assert L.mole_read_stats(None, 0, None) == M.ERR_INVALID_ARGUMENT
    L.mole_ctx_destroy(None)  # no-op


def test_parameter_update_matches_oracle_formulas():
    # host-side math only: SD and SR from sums, both SR modes, versus numpy
    rng = np.random.default_rng(0)
    n, npar = 4000, 3
    O = rng.normal(size=(n, npar)) * [1.0, 0.3, 0.1] + [0.5, -0.2, 0.05]
    E = -1.0 + 0.3 * rng.normal(size=n) + 0.1 * O[:, 0]
    s = M.SrSums()
    s.n_params = npar
    s.n = n
    s.sum_e = E.sum()
    for i in range(npar):
        s.sum_o[i] = O[:, i].sum()
        s.sum_oe[i] = (O[:, i] * E).sum()
        for j in range(npar):
            s.sum_oo[i * npar + j] = (O[:, i] * O[:, j]).sum()
    g_want = 2 * ((O * (E - E.mean())[:, None]).mean(0))
    dp, g = M.parameter_update(s, M.OPT_SD, 0.1)
    assert np.allclose(g, g_want, rtol=1e-10, atol=1e-14)
    assert np.allclose(dp, -0.1 * g_want, rtol=1e-10, atol=1e-14)
    for mode, eps in [(M.SR_TEXTBOOK, 1e-3), (M.SR_REFERENCE, 0.0)]:
        dp, g = M.parameter_update(s, M.OPT_SR, 0.05, mode, eps)
        S = O.T @ O / n
        if mode == M.SR_TEXTBOOK:
            S -= np.outer(O.mean(0), O.mean(0))
            S[np.diag_indices(npar)] *= 1 + eps
        else:
            S -= O.mean(0).sum() ** 2
            S[np.diag_indices(npar)] *= 1.01
        want = 0.05 * np.linalg.solve(S, -0.5 * g_want)
        assert np.allclose(dp, want, rtol=1e-8, atol=1e-12)
    # singular matrix -> LinalgError
    z = M.SrSums()
    z.n_params = 2
    z.n = 10
    with pytest.raises(M.MoleError) as ei:
        M.parameter_update(z, M.OPT_SR, 0.1, M.SR_TEXTBOOK, 0.0)
    assert ei.value.status == M.ERR_LINALG


def test_slater_jastrow_validation_before_device():
    from mole_amd import systems as S
    ne = S.atom(10)
    good = M.slater_jastrow(ne, 8)
    # mixing families in one context
    assert _create([good, M.helium(8)]) == M.ERR_UNSUPPORTED
    # the reference O convention is two-electron only (compat samples of
    # Slater-Jastrow systems report O = d ln psi/dp)
    assert _create([good], o_convention=M.O_REFERENCE) == M.ERR_UNSUPPORTED
    assert _create([good], precision=7) == M.ERR_INVALID_ARGUMENT
    bad = M.slater_jastrow(ne, 8)
    bad.desc.n_params = 3
    assert _create([bad]) == M.ERR_SHAPE
    bad = M.slater_jastrow(ne, 8)
    bad.desc.n_orbitals = 4
    assert _create([bad]) == M.ERR_SHAPE
    bad = M.slater_jastrow(ne, 8)
    bad.center[0] = 5
    assert _create([bad]) == M.ERR_SHAPE
    big = M.slater_jastrow(S.random_molecule(44, seed=1), 8)
    assert _create([big]) == M.ERR_SHAPE
