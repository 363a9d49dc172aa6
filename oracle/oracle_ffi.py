"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

WF_HELIUM_PADE = 1
WF_LCAO2_PADE = 2
H_ELECTRONIC = 0
H_IONIC = 1
O_LOGDERIV = 0
O_REFERENCE = 1
SR_TEXTBOOK = 0
SR_REFERENCE = 1
OPT_SD = 0
OPT_SR = 1
OPT_MOMENTUM = 2
OPT_NESTEROV = 3
OPT_LBFGS = 4
MAX_HISTORY = 64
MAX_NUC = 8


class Wf2(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("hamiltonian", C.c_int),
        ("n_nuc", C.c_int),
        ("nuc", (C.c_double * 3) * MAX_NUC),
        ("charge", C.c_int * MAX_NUC),
        ("p", C.c_double * 3),
    ]


class RunParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("sys", C.c_uint32),
        ("epoch", C.c_uint32),
        ("box_side", C.c_double),
        ("steps", C.c_uint32),
        ("block_size", C.c_uint32),
        ("init_from_stream", C.c_int),
        ("pos0", C.c_uint64),
        ("o_convention", C.c_int),
        ("all_observables", C.c_int),
    ]


class RunOut(C.Structure):
    _fields_ = [
        ("cfg", C.POINTER(C.c_double)),
        ("accept_mask", C.POINTER(C.c_uint64)),
        ("e_local", C.POINTER(C.c_double)),
        ("psi", C.POINTER(C.c_double)),
        ("pgrad", C.POINTER(C.c_double)),
        ("acceptance", C.POINTER(C.c_double)),
        ("sum_e", C.c_double),
        ("sum_e2", C.c_double),
        ("sum_bm", C.c_double),
        ("sum_bm2", C.c_double),
        ("sum_g", C.c_double * 3),
        ("n_samples", C.c_uint64),
        ("n_blocks", C.c_uint64),
        ("n_accepted", C.c_uint64),
        ("n_moves", C.c_uint64),
    ]


class OptParams(C.Structure):
    _fields_ = [
        ("optimizer", C.c_int),
        ("step", C.c_double),
        ("sr_mode", C.c_int),
        ("sr_eps", C.c_double),
        ("o_convention", C.c_int),
        ("persistent", C.c_int),
    ]


class StatefulOpt(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("np", C.c_int),
        ("step", C.c_double),
        ("mu", C.c_double),
        ("history", C.c_int),
        ("iter", C.c_size_t),
        ("momentum", C.c_double * 16),
        ("momentum_prev", C.c_double * 16),
        ("grad_prev", C.c_double * 16),
        ("pars_prev", C.c_double * 16),
        ("len", C.c_int),
        ("s", (C.c_double * 16) * MAX_HISTORY),
        ("y", (C.c_double * 16) * MAX_HISTORY),
    ]


_lib = None


def lib():
    """Load the oracle library (build it with `make -C oracle`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    L = C.CDLL(LIB_PATH)
    d, u32, u64, i = C.c_double, C.c_uint32, C.c_uint64, C.c_int
    pd = C.POINTER(C.c_double)
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.orc_stream_u64.restype = u64
    L.orc_stream_u64.argtypes = [u64, u32, u32, u32, u64]
    L.orc_u64_standard.restype = d
    L.orc_u64_standard.argtypes = [u64]
    L.orc_u64_range.restype = d
    L.orc_u64_range.argtypes = [u64, d, d]
    L.orc_exp.restype = d
    L.orc_exp.argtypes = [d]
    L.orc_norm3.restype = d
    L.orc_norm3.argtypes = [d, d, d]
    L.orc_log.restype = d
    L.orc_log.argtypes = [d]
    L.orc_sincospi.argtypes = [d, pd, pd]
    L.orc_box_muller.argtypes = [u64, u64, pd, pd]
    pw = C.POINTER(Wf2)
    L.orc_wf2_value.restype = d
    L.orc_wf2_value.argtypes = [pw, pd]
    L.orc_wf2_gradient.argtypes = [pw, pd, pd]
    L.orc_wf2_laplacian.restype = d
    L.orc_wf2_laplacian.argtypes = [pw, pd]
    L.orc_wf2_parameter_gradient.argtypes = [pw, pd, pd]
    L.orc_wf2_local_energy.restype = d
    L.orc_wf2_local_energy.argtypes = [pw, pd]
    L.orc_ionic_potential.restype = d
    L.orc_ionic_potential.argtypes = [pw, pd, i]
    L.orc_electronic_potential.restype = d
    L.orc_electronic_potential.argtypes = [pd, i]
    L.orc_run_box.restype = i
    L.orc_run_box.argtypes = [pw, C.POINTER(RunParams), u32, u32, i, C.POINTER(RunOut)]
    L.orc_running_mean.restype = d
    L.orc_running_mean.argtypes = [pd, C.c_size_t]
    L.orc_blocking_error.restype = d
    L.orc_blocking_error.argtypes = [pd, C.c_size_t, C.c_size_t]
    L.orc_energy_gradient.argtypes = [pd, pd, pd, C.c_size_t, i, d, pd]
    L.orc_sr_matrix.argtypes = [pd, pd, C.c_size_t, i, i, d, pd]
    L.orc_solve_sym.restype = i
    L.orc_solve_sym.argtypes = [pd, pd, i, pd]
    L.orc_vmc_optimize.restype = i
    L.orc_vmc_optimize.argtypes = [pw, u64, d, u32, u32, u32, u32, C.POINTER(OptParams), i,
                                   pd, pd, pd, pd, pd]
    L.orc_stateful_opt_init.restype = i
    L.orc_stateful_opt_init.argtypes = [C.POINTER(StatefulOpt), i, i, d, d, i]
    L.orc_stateful_opt_update.argtypes = [C.POINTER(StatefulOpt), pd, pd, pd]
    L.orc_tau_corr.restype = d
    L.orc_tau_corr.argtypes = [pd, C.c_size_t]
    _lib = L
    return L


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def u64ptr(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def make_wf2(kind, params, nuclei, charges, hamiltonian=H_ELECTRONIC):
    wf = Wf2()
    wf.kind = kind
    wf.hamiltonian = hamiltonian
    wf.n_nuc = len(nuclei)
    for I, (R, Z) in enumerate(zip(nuclei, charges)):
        for k in range(3):
            wf.nuc[I][k] = float(R[k])
        wf.charge[I] = int(Z)
    for k in range(3):
        wf.p[k] = float(params[k])
    return wf


def helium(zeta=1.6875, b=0.5, c=0.3, hamiltonian=H_ELECTRONIC):
    return make_wf2(WF_HELIUM_PADE, (zeta, b, c), [(0.0, 0.0, 0.0)], [2], hamiltonian)


def h2(zeta=1.0, b=0.5, c=0.3, half_bond=0.7):
    return make_wf2(WF_LCAO2_PADE, (zeta, b, c),
                    [(0.0, 0.0, -half_bond), (0.0, 0.0, half_bond)], [1, 1])


def log(x):
    return lib().orc_log(float(x))


def sincospi(x):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincospi(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def box_muller(a, b):
    z0, z1 = C.c_double(), C.c_double()
    lib().orc_box_muller(int(a), int(b), C.byref(z0), C.byref(z1))
    return z0.value, z1.value


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return tuple(o)


def run_box(wf, *, seed, walkers, steps, block_size, box_side, epoch=0, sys=0,
            cfg=None, pos0=0, o_convention=O_LOGDERIV, n_threads=1,
            record=True, pgrad=False, all_observables=False):
    """Run `walkers` independent Sampler+Runner::run chains on the oracle."""
    L = lib()
    rp = RunParams(seed=seed, sys=sys, epoch=epoch, box_side=box_side, steps=steps,
                   block_size=block_size, init_from_stream=0 if cfg is not None else 1,
                   pos0=pos0, o_convention=o_convention, all_observables=int(all_observables))
    blocks = steps // block_size
    ns = (blocks - 1) * block_size
    res = {}
    out = RunOut()
    cfg_arr = np.zeros((walkers, 6)) if cfg is None else np.ascontiguousarray(cfg, dtype=np.float64).copy()
    out.cfg = dptr(cfg_arr)
    acc = np.zeros(walkers)
    out.acceptance = dptr(acc)
    if record:
        masks = np.zeros((walkers, blocks * block_size), dtype=np.uint64)
        el = np.zeros((walkers, ns))
        psi = np.zeros((walkers, ns))
        out.accept_mask = u64ptr(masks)
        out.e_local = dptr(el)
        out.psi = dptr(psi)
        res.update(accept_mask=masks, e_local=el, psi=psi)
        if pgrad:
            pg = np.zeros((walkers, ns, 3))
            out.pgrad = dptr(pg)
            res["pgrad"] = pg
    rc = L.orc_run_box(C.byref(wf), C.byref(rp), 0, walkers, n_threads, C.byref(out))
    if rc != 0:
        raise ValueError("orc_run_box: steps < 2*block_size")
    res.update(cfg=cfg_arr, acceptance=acc, sum_e=out.sum_e, sum_e2=out.sum_e2,
               sum_bm=out.sum_bm, sum_bm2=out.sum_bm2, n_samples=out.n_samples,
               n_blocks=out.n_blocks, n_accepted=out.n_accepted, n_moves=out.n_moves)
    return res


class OracleOptimizer:
    """orc_stateful_opt_*: MomentumDescent / NesterovMomentum / OnlineLbfgs on a given gradient."""

    def __init__(self, kind, np_, step, mu=0.0, history=1):
        self.st = StatefulOpt()
        assert lib().orc_stateful_opt_init(C.byref(self.st), kind, np_, step, mu, history) == 0
        self.np = np_

    def update(self, pars, g):
        pars = np.ascontiguousarray(pars, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        dp = np.zeros(self.np)
        lib().orc_stateful_opt_update(C.byref(self.st), dptr(pars), dptr(g), dptr(dp))
        return dp
