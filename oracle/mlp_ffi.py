"""ctypes access to the oracle's learned-proposal restatement (oracle/mlp.c).
Test infrastructure: used by tests/ and by the bench's CPU baseline only."""
import ctypes as C

import numpy as np

from oracle import oracle_ffi as OF
from oracle import sj_ffi as SJ

NPARAM = 9476
NIN, HID, NOUT = 13, 64, 4
OFF = dict(W1=0, b1=832, W2=896, b2=4992, W3=5056, b3=9152, W4=9216, b4=9472)
LS0 = -1.3862943611198906
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_bound = False


def lib():
    global _bound
    L = SJ.lib()
    if not _bound:
        ps = C.POINTER(SJ.Sj)
        L.orc_mlp_features.argtypes = [ps, _pd, C.c_int, _pd, _pd]
        L.orc_bf16.restype = C.c_float
        L.orc_bf16.argtypes = [C.c_float]
        L.orc_mlp_forward.argtypes = [_pf, _pd, C.c_int, _pd, _pd]
        L.orc_mlp_log_sigma.restype = C.c_double
        L.orc_mlp_log_sigma.argtypes = [C.c_double]
        L.orc_mlp_log_sigma_slope.restype = C.c_double
        L.orc_mlp_log_sigma_slope.argtypes = [C.c_double]
        L.orc_mlp_sample.restype = C.c_double
        L.orc_mlp_sample.argtypes = [ps, _pf, _pd, C.c_int, _pd, C.c_int, _pd, _pd, _pd, C.c_int]
        L.orc_mlp_log_q_ratio.restype = C.c_double
        L.orc_mlp_log_q_ratio.argtypes = [ps, _pf, _pd, C.c_int, _pd, C.c_int]
        L.orc_sj_move_terms.argtypes = [ps, _pd, C.c_int, _pd, _pd, _pd]
        _bound = True
    return L


def log_sigma(out3):
    return lib().orc_mlp_log_sigma(float(out3))


def log_sigma_slope(out3):
    return lib().orc_mlp_log_sigma_slope(float(out3))


def _d(a):
    return a.ctypes.data_as(_pd)


def _f(a):
    return a.ctypes.data_as(_pf)


def features(wv, cfg, i, r=None):
    cfg = np.ascontiguousarray(cfg, dtype=np.float64)
    r = cfg[3 * i:3 * i + 3].copy() if r is None else np.ascontiguousarray(r, dtype=np.float64)
    f = np.zeros(NIN)
    lib().orc_mlp_features(wv.ref(), _d(cfg), int(i), _d(r), _d(f))
    return f


def forward(w, f, emulate_bf16=False):
    w = np.ascontiguousarray(w, dtype=np.float32)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(NOUT)
    act = np.zeros(3 * HID)
    lib().orc_mlp_forward(_f(w), _d(f), int(emulate_bf16), _d(out), _d(act))
    return out, act.reshape(3, HID)


OBJ_ESJD, OBJ_ESJD_RELATIVE, OBJ_LOG_ESJD = 0, 1, 2


def sample(wv, w, cfg, i, z, emulate_bf16=False, frozen_f2=None, grad=False, objective=OBJ_ESJD):
    """(O, dO/dtheta or None, diag) of one training sample."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    cfg = np.ascontiguousarray(cfg, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    g = np.zeros(NPARAM) if grad else None
    diag = np.zeros(9)
    ff = None if frozen_f2 is None else np.ascontiguousarray(frozen_f2, dtype=np.float64)
    O = lib().orc_mlp_sample(wv.ref(), _f(w), _d(cfg), int(i), _d(z), int(emulate_bf16),
                             None if ff is None else _d(ff), None if g is None else _d(g), _d(diag),
                             int(objective))
    return O, g, dict(logr=diag[0], A=diag[1], d2=diag[2], ls=diag[3], ls2=diag[4], lR2=diag[5],
                      xn=diag[6:9].copy())


def log_q_ratio(wv, w, cfg, i, xn, emulate_bf16=False):
    w = np.ascontiguousarray(w, dtype=np.float32)
    cfg = np.ascontiguousarray(cfg, dtype=np.float64)
    xn = np.ascontiguousarray(xn, dtype=np.float64)
    return lib().orc_mlp_log_q_ratio(wv.ref(), _f(w), _d(cfg), int(i), _d(xn), int(emulate_bf16))


def move_terms(wv, cfg, i, xn):
    cfg = np.ascontiguousarray(cfg, dtype=np.float64)
    xn = np.ascontiguousarray(xn, dtype=np.float64)
    lR2 = C.c_double()
    g = np.zeros(3)
    lib().orc_sj_move_terms(wv.ref(), _d(cfg), int(i), _d(xn), C.byref(lR2), _d(g))
    return lR2.value, g


def xavier(seed=2):
    """Host Xavier-uniform weights for tests (numpy generator; the device
    initialiser has its own Philox stream)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, np.float32)
    for name, (nin, nout) in dict(W1=(NIN, HID), W2=(HID, HID), W3=(HID, HID), W4=(HID, NOUT)).items():
        a = np.sqrt(6.0 / (nin + nout))
        w[OFF[name]:OFF[name] + nin * nout] = rng.uniform(-a, a, nin * nout)
    return w
