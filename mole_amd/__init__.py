"""Python access to libmole_amd.so, the gfx950 walker hot path of mole.

Thin ctypes plumbing over the C ABI in include/mole_gpu.h: no computation
happens in Python and there is no CPU fallback -- if the library is missing
or no device is present, calls fail with MoleError.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmole_amd.so")

# status codes (include/mole_gpu.h)
OK = 0
ERR_LINALG = 1
ERR_SHAPE = 2
ERR_FUNC = 3
ERR_OPERATOR_VALUE_ACCESS = 4
ERR_DATA_ACCESS = 5
ERR_EMPTY_CACHE = 6
ERR_INVALID_ARGUMENT = 16
ERR_DEVICE = 17
ERR_UNSUPPORTED = 18
ERR_OUT_OF_MEMORY = 19

WF_HELIUM_PADE = 1
WF_LCAO2_PADE = 2
WF_SLATER_JASTROW = 3
PREC_FP64 = 0
PREC_MIXED = 1
H_ELECTRONIC = 0
H_IONIC = 1
PROPOSAL_BOX = 0
PROPOSAL_DIFFUSE = 1
PROPOSAL_LEARNED = 2
MLP_PARAMS = 9476
TAU_NOALLOC_MAX_WALKERS = 262144  # mole_tau_int_noalloc's scratch (MOLE_TAU_NOALLOC_MAX_WALKERS)
O_LOGDERIV = 0
O_REFERENCE = 1
FLAG_TRACE = 1
FLAG_SAMPLES = 2
SWEEP_SAMPLE = 1
VARIANT_AUTO = 0
VARIANT_LANE = 1
VARIANT_WAVESPEC = 2
VARIANT_QUAD = 3
SCHEDULE_CLASSES, SCHEDULE_PERSISTENT = 0, 1
MAX_PARAMS = 16

BASIS_STO = 0
BASIS_GTO = 1
BASIS_STO_PX, BASIS_STO_PY, BASIS_STO_PZ = 2, 3, 4
TERM_KINETIC = 0
TERM_IONIC_POTENTIAL = 1
TERM_ELECTRONIC_POTENTIAL = 2
TERM_HARMONIC = 3
TERM_RADIAL_POWER = 4
TERM_DIPOLE = 5
OPERATOR_MAX_TERMS = 4
MAX_OBSERVABLES = 8
OPT_SD = 0
OPT_SR = 1
OPT_MOMENTUM = 2
BRANCH_SR = 0
BRANCH_SIMPLE = 1
OPT_NESTEROV = 3
OPT_LBFGS = 4
LBFGS_MAX_HISTORY = 16
SR_TEXTBOOK = 0
SR_REFERENCE = 1

DEBUG_EXP = 0
DEBUG_SQRT = 1
DEBUG_DIV = 2
DEBUG_STANDARD = 3
DEBUG_RANGE = 4
DEBUG_NORM3 = 5
DEBUG_LOG = 6
DEBUG_SINPI = 7
DEBUG_COSPI = 8
DEBUG_BOXMULLER0 = 9
DEBUG_BOXMULLER1 = 10
DEBUG_BF16 = 11


class MoleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


class TrainDesc(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64), ("objective", C.c_uint32)]


MLP_OBJ_ESJD, MLP_OBJ_ESJD_RELATIVE, MLP_OBJ_LOG_ESJD = 0, 1, 2


class SystemDesc(C.Structure):
    _fields_ = [
        ("wf_kind", C.c_uint32),
        ("hamiltonian", C.c_uint32),
        ("n_up", C.c_uint32),
        ("n_down", C.c_uint32),
        ("n_nuclei", C.c_uint32),
        ("nuclei_xyz", C.POINTER(C.c_double)),
        ("nuclei_charge", C.POINTER(C.c_int32)),
        ("n_params", C.c_uint32),
        ("params", C.POINTER(C.c_double)),
        ("n_walkers", C.c_uint32),
        ("n_basis", C.c_uint32),
        ("basis_exponent", C.POINTER(C.c_double)),
        ("basis_center", C.POINTER(C.c_uint32)),
        ("n_orbitals", C.c_uint32),
        ("orbital_coeff", C.POINTER(C.c_double)),
        ("basis_kind", C.POINTER(C.c_uint32)),
    ]


class RunDesc(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("proposal", C.c_uint32),
        ("step", C.c_double),
        ("block_size", C.c_uint32),
        ("o_convention", C.c_uint32),
        ("flags", C.c_uint32),
        ("record_sweeps", C.c_uint32),
        ("precision", C.c_uint32),
        ("lu_period", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_samples", C.c_uint64),
        ("n_blocks", C.c_uint64),
        ("n_moves", C.c_uint64),
        ("n_accepted", C.c_uint64),
        ("sum_e", C.c_double),
        ("sum_e2", C.c_double),
        ("sum_block_mean", C.c_double),
        ("sum_block_mean2", C.c_double),
        ("max_condition", C.c_double),
        ("n_singular", C.c_uint64),
        ("max_logdet_drift", C.c_double),
    ]


class SrSums(C.Structure):
    _fields_ = [
        ("n_params", C.c_uint32),
        ("n", C.c_uint64),
        ("sum_e", C.c_double),
        ("sum_o", C.c_double * MAX_PARAMS),
        ("sum_oe", C.c_double * MAX_PARAMS),
        ("sum_oo", C.c_double * (MAX_PARAMS * MAX_PARAMS)),
    ]


class Term(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("coef", C.c_double), ("param", C.c_double * 6)]


class Operator(C.Structure):
    """mole_operator: sum of coef * term (see TERM_*)."""
    _fields_ = [("n_terms", C.c_uint32), ("term", Term * OPERATOR_MAX_TERMS)]

    @classmethod
    def of(cls, terms):
        """[(kind, coef, *params)] -> Operator."""
        op = cls(n_terms=len(terms))
        for k, t in enumerate(terms):
            op.term[k].kind = int(t[0])
            op.term[k].coef = float(t[1])
            for q, v in enumerate(t[2:]):
                op.term[k].param[q] = float(v)
        return op


def harmonic_hamiltonian(omega):
    """HarmonicHamiltonian of examples/custom_operator.rs:30-62: T + 1/2 w^2 |x|^2."""
    return [(TERM_KINETIC, 1.0), (TERM_HARMONIC, 1.0, omega, 0.0, 0.0, 0.0)]


def electronic_hamiltonian():
    """ElectronicHamiltonian (operator.rs:151-183) as terms."""
    return [(TERM_KINETIC, 1.0), (TERM_IONIC_POTENTIAL, 1.0), (TERM_ELECTRONIC_POTENTIAL, 1.0)]


class BlockRecord(C.Structure):
    _fields_ = [
        ("block", C.c_uint64),
        ("n_walkers", C.c_uint64),
        ("mean", C.c_double),
        ("sum_block_mean", C.c_double),
        ("sum_block_mean2", C.c_double),
    ]


class OptimizerState(C.Structure):
    """mole_optimizer: the state of a stateful optimizer (plain data)."""
    _fields_ = [
        ("kind", C.c_uint32),
        ("n_params", C.c_uint32),
        ("sr_mode", C.c_uint32),
        ("history", C.c_uint32),
        ("step", C.c_double),
        ("sr_eps", C.c_double),
        ("momentum", C.c_double),
        ("iter", C.c_uint64),
        ("n_pairs", C.c_uint32),
        ("head", C.c_uint32),
        ("m", C.c_double * MAX_PARAMS),
        ("m_prev", C.c_double * MAX_PARAMS),
        ("grad_prev", C.c_double * MAX_PARAMS),
        ("pars_prev", C.c_double * MAX_PARAMS),
        ("s", (C.c_double * MAX_PARAMS) * LBFGS_MAX_HISTORY),
        ("y", (C.c_double * MAX_PARAMS) * LBFGS_MAX_HISTORY),
    ]


_lib = None

# (name, restype, argtypes)
_P = C.c_void_p
_pd = C.POINTER(C.c_double)
_pu64 = C.POINTER(C.c_uint64)
_pf = C.POINTER(C.c_float)
_SIGS = [
    ("mole_status_string", C.c_char_p, [C.c_int32]),
    ("mole_last_error", C.c_char_p, []),
    ("mole_device_info", C.c_int32, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int32)]),
    ("mole_ctx_create", C.c_int32, [C.POINTER(SystemDesc), C.c_uint32, C.POINTER(RunDesc), C.POINTER(_P)]),
    ("mole_ctx_destroy", None, [_P]),
    ("mole_init_walkers", C.c_int32, [_P]),
    ("mole_set_configs", C.c_int32, [_P, C.c_uint32, _pd, C.c_size_t]),
    ("mole_get_configs", C.c_int32, [_P, C.c_uint32, _pd, C.c_size_t]),
    ("mole_broadcast_config", C.c_int32, [_P, C.c_uint32, _pd]),
    ("mole_set_stream", C.c_int32, [_P, C.c_uint32, C.c_uint64]),
    ("mole_get_stream", C.c_int32, [_P, C.POINTER(C.c_uint32), _pu64]),
    ("mole_set_params", C.c_int32, [_P, C.c_uint32, _pd, C.c_uint32]),
    ("mole_get_params", C.c_int32, [_P, C.c_uint32, _pd, C.c_uint32]),
    ("mole_sweep", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("mole_run", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("mole_reset_stats", C.c_int32, [_P]),
    ("mole_read_stats", C.c_int32, [_P, C.c_uint32, C.POINTER(Stats)]),
    ("mole_read_sr_sums", C.c_int32, [_P, C.c_uint32, C.POINTER(SrSums)]),
    ("mole_read_trace", C.c_int32, [_P, C.c_uint32, _pu64, _pd, C.c_size_t, C.c_size_t, C.c_size_t]),
    ("mole_read_samples", C.c_int32, [_P, C.c_uint32, _pd, _pd, _pd, C.c_size_t, C.c_size_t]),
    ("mole_read_logdet", C.c_int32, [_P, C.c_uint32, _pd, _pd, C.c_size_t]),
    ("mole_tau_int", C.c_int32, [_P, C.c_uint32, C.c_size_t, C.c_size_t, _pd, C.c_size_t]),
    ("mole_tau_int_noalloc", C.c_int32, [_P, C.c_uint32, C.c_size_t, C.c_size_t, _pd, C.c_size_t]),
    ("mole_eval", C.c_int32, [_P, C.c_uint32, C.c_size_t, _pd, _pd, _pd, _pd, _pd, _pd]),
    ("mole_set_launch_config", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("mole_set_walker_lanes", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("mole_get_walker_lanes", C.c_int32, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("mole_get_launch_info", C.c_int32, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_size_t)]),
    ("mole_set_tile_schedule", C.c_int32, [_P, C.c_uint32, C.c_uint32]),
    ("mole_graph_capture", C.c_int32, [_P, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    ("mole_graph_launch", C.c_int32, [_P, _P, C.c_uint32]),
    ("mole_graph_destroy", None, [_P]),
    ("mole_synchronize", C.c_int32, [_P]),
    ("mole_last_timing", C.c_int32, [_P, _pd, C.POINTER(C.c_uint32)]),
    ("mole_compute_parameter_update", C.c_int32, [C.POINTER(SrSums), C.c_uint32, C.c_double, C.c_uint32, C.c_double, _pd, _pd]),
    ("mole_optimizer_init", C.c_int32, [C.POINTER(OptimizerState), C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                        C.c_uint32, C.c_uint32, C.c_double]),
    ("mole_optimizer_update", C.c_int32, [C.POINTER(OptimizerState), C.POINTER(SrSums), _pd, _pd, _pd]),
    ("mole_set_hamiltonian", C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(Operator)]),
    ("mole_set_observables", C.c_int32, [C.c_void_p, C.c_uint32, C.POINTER(Operator), C.c_uint32]),
    ("mole_read_observables", C.c_int32, [C.c_void_p, C.c_uint32, _pd, _pd, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("mole_read_blocks", C.c_int32, [C.c_void_p, C.c_uint32, C.c_size_t, C.POINTER(BlockRecord), C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    ("mole_dmc_init", C.c_int32, [C.c_void_p, _pd, C.c_uint32]),
    ("mole_dmc_diffuse", C.c_int32, [C.c_void_p, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("mole_dmc_read", C.c_int32, [C.c_void_p, C.c_uint32, _pd, _pd, C.c_size_t, C.POINTER(C.c_size_t), _pd]),
    ("mole_dmc_read_steps", C.c_int32, [C.c_void_p, C.c_uint32, _pd, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("mole_dmc_read_walkers", C.c_int32, [C.c_void_p, C.c_uint32, _pd, _pd, C.c_size_t]),
    ("mole_optimize_with", C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(OptimizerState),
                                       _pd, _pd]),
    ("mole_debug_philox", C.c_int32, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_size_t, _pu64]),
    ("mole_debug_math", C.c_int32, [C.c_int32, C.c_size_t, _pd, _pd, _pd]),
    ("mole_debug_copy", C.c_int32, [C.c_size_t, C.c_int32, _pd]),
    ("mole_debug_mfma_bf16", C.c_int32, [_pf, _pf, _pf]),
    ("mole_debug_mfma_bf16_k32", C.c_int32, [_pf, _pf, _pf]),
    ("mole_debug_mlp_features", C.c_int32, [_pd, C.c_uint32, _pd, C.c_uint32, C.POINTER(C.c_int32), C.c_size_t,
                                            _pf]),
    ("mole_optimize", C.c_int32, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32,
                                  C.c_double, _pd, _pd]),
    ("mole_mlp_create", C.c_int32, [C.c_uint64, C.POINTER(_P)]),
    ("mole_mlp_destroy", None, [_P]),
    ("mole_mlp_set_weights", C.c_int32, [_P, _pf, C.c_size_t]),
    ("mole_mlp_get_weights", C.c_int32, [_P, _pf, C.c_size_t]),
    ("mole_mlp_attach", C.c_int32, [_P, _P]),
    ("mole_mlp_collect", C.c_int32, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("mole_mlp_clear_data", C.c_int32, [_P]),
    ("mole_mlp_data_size", C.c_int32, [_P, C.POINTER(C.c_size_t)]),
    ("mole_mlp_train", C.c_int32, [_P, _P, C.c_uint32, C.POINTER(TrainDesc), _pd]),
    ("mole_mlp_gradient", C.c_int32, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, _pf, _pd]),
    ("mole_mlp_eval", C.c_int32, [_P, _pd, C.c_uint32, _pd, C.c_uint32, C.POINTER(C.c_int32), C.c_uint32, _pf]),
]


def lib():
    """Load libmole_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MoleError(ERR_DEVICE, f"{LIB_PATH} is missing: build it with python -m mole_amd.build")
    L = C.CDLL(LIB_PATH)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return [s[0] for s in _SIGS]


def check(status):
    if status != OK:
        L = lib()
        raise MoleError(status, f"{L.mole_status_string(status).decode()}: {L.mole_last_error().decode()}")


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pd)


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_pf)


def _u64p(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pu64)


class System:
    """Host description of one system (keeps the arrays the descriptor points into)."""

    def __init__(self, wf_kind, nuclei, charges, params, n_walkers, n_up=1, n_down=1,
                 hamiltonian=H_ELECTRONIC):
        self.nuclei = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 3)
        self.charges = np.ascontiguousarray(charges, dtype=np.int32)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.desc = SystemDesc(
            wf_kind=wf_kind, hamiltonian=hamiltonian, n_up=n_up, n_down=n_down,
            n_nuclei=self.nuclei.shape[0],
            nuclei_xyz=self.nuclei.ctypes.data_as(_pd),
            nuclei_charge=self.charges.ctypes.data_as(C.POINTER(C.c_int32)),
            n_params=self.params.size, params=self.params.ctypes.data_as(_pd),
            n_walkers=n_walkers)
        self.n_electrons = n_up + n_down
        self.n_walkers = n_walkers

    def with_basis(self, zeta, center, coef, kind=None):
        """Attach the s-type basis (STO, or GTO where kind[m] = BASIS_GTO) and the
        orbital coefficients (Slater-Jastrow)."""
        if kind is not None:
            self.kind = np.ascontiguousarray(kind, dtype=np.uint32)
            self.desc.basis_kind = self.kind.ctypes.data_as(C.POINTER(C.c_uint32))
        self.zeta = np.ascontiguousarray(zeta, dtype=np.float64)
        self.center = np.ascontiguousarray(center, dtype=np.uint32)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1, self.zeta.size)
        self.desc.n_basis = self.zeta.size
        self.desc.basis_exponent = self.zeta.ctypes.data_as(_pd)
        self.desc.basis_center = self.center.ctypes.data_as(C.POINTER(C.c_uint32))
        self.desc.n_orbitals = self.coef.shape[0]
        self.desc.orbital_coeff = self.coef.ctypes.data_as(_pd)
        return self


def slater_jastrow(sj, n_walkers, hamiltonian=H_ELECTRONIC, params=None, exponent_scale=None):
    """Device system from a mole_amd.systems.SJSystem.  params: the five
    Jastrow parameters, or six with the exponent scale s (basis exponents s
    zeta_m) as a variational parameter; exponent_scale appends s to the
    system's Jastrow parameters."""
    p = sj.jastrow if params is None else params
    if exponent_scale is not None:
        p = list(p)[:5] + [float(exponent_scale)]
    s = System(WF_SLATER_JASTROW, sj.nuclei, sj.charges, p, n_walkers, n_up=sj.n_up,
               n_down=sj.n_dn, hamiltonian=hamiltonian)
    return s.with_basis(sj.zeta, sj.center, sj.coef, getattr(sj, "kind", None))


def helium(n_walkers, zeta=1.6875, b=0.5, c=0.3, hamiltonian=H_ELECTRONIC):
    return System(WF_HELIUM_PADE, [(0.0, 0.0, 0.0)], [2], [zeta, b, c], n_walkers,
                  hamiltonian=hamiltonian)


def h2(n_walkers, zeta=1.0, b=0.5, c=0.3, half_bond=0.7):
    return System(WF_LCAO2_PADE, [(0.0, 0.0, -half_bond), (0.0, 0.0, half_bond)], [1, 1],
                  [zeta, b, c], n_walkers)


class Context:
    """One mole_ctx: a batch of walkers of one or more systems on the device."""

    def __init__(self, systems, *, seed=0, step=0.5, block_size=100, o_convention=O_LOGDERIV,
                 flags=0, record_sweeps=0, proposal=PROPOSAL_BOX, precision=PREC_FP64,
                 lu_period=0):
        L = lib()
        self.systems = list(systems)
        arr = (SystemDesc * len(self.systems))(*[s.desc for s in self.systems])
        self.run_desc = RunDesc(seed=seed, proposal=proposal, step=step, block_size=block_size,
                                o_convention=o_convention, flags=flags,
                                record_sweeps=record_sweeps, precision=precision,
                                lu_period=lu_period)
        h = C.c_void_p()
        check(L.mole_ctx_create(arr, len(self.systems), C.byref(self.run_desc), C.byref(h)))
        self.h = h
        self.block_size = block_size
        self.record_sweeps = record_sweeps

    def close(self):
        if getattr(self, "h", None):
            lib().mole_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_walkers(self):
        check(lib().mole_init_walkers(self.h))

    def set_configs(self, cfg, sys=0):
        s = self.systems[sys]
        a = np.ascontiguousarray(cfg, dtype=np.float64).reshape(s.n_walkers, s.n_electrons * 3)
        check(lib().mole_set_configs(self.h, sys, _dp(a), s.n_walkers))

    def get_configs(self, sys=0):
        s = self.systems[sys]
        a = np.zeros((s.n_walkers, s.n_electrons * 3))
        check(lib().mole_get_configs(self.h, sys, _dp(a), s.n_walkers))
        return a

    def broadcast_config(self, cfg, sys=0):
        a = np.ascontiguousarray(cfg, dtype=np.float64).ravel()
        check(lib().mole_broadcast_config(self.h, sys, _dp(a)))

    def set_stream(self, epoch, position):
        check(lib().mole_set_stream(self.h, epoch, position))

    def get_stream(self):
        e, p = C.c_uint32(), C.c_uint64()
        check(lib().mole_get_stream(self.h, C.byref(e), C.byref(p)))
        return e.value, p.value

    def set_params(self, p, sys=0):
        a = np.ascontiguousarray(p, dtype=np.float64)
        check(lib().mole_set_params(self.h, sys, _dp(a), a.size))

    def get_params(self, sys=0):
        a = np.zeros(self.systems[sys].params.size)
        check(lib().mole_get_params(self.h, sys, _dp(a), a.size))
        return a

    def sweep(self, n, sample=False):
        check(lib().mole_sweep(self.h, n, SWEEP_SAMPLE if sample else 0))

    def run(self, steps, block_size=None):
        check(lib().mole_run(self.h, steps, self.block_size if block_size is None else block_size))

    def reset_stats(self):
        check(lib().mole_reset_stats(self.h))

    def stats(self, sys=0):
        s = Stats()
        check(lib().mole_read_stats(self.h, sys, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def sr_sums_raw(self, sys=0):
        s = SrSums()
        check(lib().mole_read_sr_sums(self.h, sys, C.byref(s)))
        return s

    def sr_sums(self, sys=0):
        s = self.sr_sums_raw(sys)
        np_ = s.n_params
        return dict(n=s.n, sum_e=s.sum_e, sum_o=np.array(s.sum_o[:np_]),
                    sum_oe=np.array(s.sum_oe[:np_]),
                    sum_oo=np.array(s.sum_oo[:np_ * np_]).reshape(np_, np_))

    def trace(self, sweeps, samples, sys=0):
        W = self.systems[sys].n_walkers
        m = np.zeros((W, sweeps), dtype=np.uint64)
        e = np.zeros((W, samples))
        check(lib().mole_read_trace(self.h, sys, _u64p(m), _dp(e), W, sweeps, samples))
        return m, e

    def logdet(self, sys=0):
        """mole_read_logdet: (log|det A_s| [W][2], sign [W][2]) per walker."""
        W = self.systems[sys].n_walkers
        ld = np.zeros((W, 2))
        sg = np.zeros((W, 2))
        check(lib().mole_read_logdet(self.h, sys, _dp(ld), _dp(sg), W))
        return ld, sg

    def tau_int(self, n, first=0, sys=0):
        """mole_tau_int: per-walker integrated autocorrelation time of the
        local-energy samples [first, first + n) of the trace, on the device."""
        W = self.systems[sys].n_walkers
        t = np.zeros(W)
        check(lib().mole_tau_int(self.h, sys, first, n, _dp(t), W))
        return t

    def tau_int_noalloc(self, n, first=0, sys=0):
        """mole_tau_int_noalloc: as tau_int, staged in the library's static
        device scratch instead of an allocation."""
        W = self.systems[sys].n_walkers
        t = np.zeros(W)
        check(lib().mole_tau_int_noalloc(self.h, sys, first, n, _dp(t), W))
        return t

    def samples(self, samples, sys=0):
        W = self.systems[sys].n_walkers
        npar = self.systems[sys].params.size
        e = np.zeros((W, samples))
        p = np.zeros((W, samples))
        g = np.zeros((W, samples, npar))
        check(lib().mole_read_samples(self.h, sys, _dp(e), _dp(p), _dp(g), W, samples))
        return e, p, g

    def eval(self, cfg, sys=0):
        s = self.systems[sys]
        D = 3 * s.n_electrons
        a = np.ascontiguousarray(cfg, dtype=np.float64).reshape(-1, D)
        n = a.shape[0]
        out = dict(value=np.zeros(n), grad=np.zeros((n, D)), lap=np.zeros(n),
                   e_local=np.zeros(n), pgrad=np.zeros((n, s.params.size)))
        check(lib().mole_eval(self.h, sys, n, _dp(a), _dp(out["value"]), _dp(out["grad"]),
                              _dp(out["lap"]), _dp(out["e_local"]), _dp(out["pgrad"])))
        return out

    def optimize(self, iters, steps, block_size, optimizer=None, step=0.05, sr_mode=None, sr_eps=1e-3):
        """mole_optimize: iters optimizer iterations with the loop on the device.
        Returns (params, log[iters][8])."""
        optimizer = OPT_SR if optimizer is None else optimizer
        sr_mode = SR_TEXTBOOK if sr_mode is None else sr_mode
        log = np.zeros((iters, 8))
        p = np.zeros(self.systems[0].params.size)
        check(lib().mole_optimize(self.h, iters, steps, block_size, optimizer, step, sr_mode, sr_eps, _dp(log),
                                  _dp(p)))
        return p, log

    def optimize_with(self, opt, iters, steps, block_size):
        """mole_optimize_with: the device loop with any Optimizer (its state
        advances in place).  Returns (params, log[iters][8])."""
        log = np.zeros((iters, 8))
        p = np.zeros(self.systems[0].params.size)
        check(lib().mole_optimize_with(self.h, iters, steps, block_size, C.byref(opt.state), _dp(log), _dp(p)))
        return p, log

    # ---- local operators on the device ----

    def set_hamiltonian(self, terms, sys=0):
        """mole_set_hamiltonian: the "Energy" operator as [(kind, coef, *params)]; None restores
        the built-in Hamiltonian."""
        if terms is None:
            check(lib().mole_set_hamiltonian(self.h, sys, None))
        else:
            op = Operator.of(terms)
            check(lib().mole_set_hamiltonian(self.h, sys, C.byref(op)))

    def set_observables(self, operators, sys=0):
        """mole_set_observables: a list of operators (each a list of terms)."""
        n = len(operators)
        arr = (Operator * max(1, n))(*[Operator.of(t) for t in operators])
        check(lib().mole_set_observables(self.h, sys, arr if n else None, n))
        self._n_obs = getattr(self, "_n_obs", {})
        self._n_obs[sys] = n

    def read_observables(self, sys=0):
        """(sum, sum of squares, n_samples) of the observables."""
        n = getattr(self, "_n_obs", {}).get(sys, 0)
        s1, s2 = np.zeros(max(1, n)), np.zeros(max(1, n))
        ns = C.c_uint64()
        check(lib().mole_read_observables(self.h, sys, _dp(s1), _dp(s2), n, C.byref(ns)))
        return s1[:n], s2[:n], ns.value

    def read_blocks(self, first=0, sys=0):
        """mole_read_blocks: per-block energy records from block `first` on,
        as a structured array (block, n_walkers, mean, sum_block_mean,
        sum_block_mean2)."""
        n = C.c_size_t()
        check(lib().mole_read_blocks(self.h, sys, first, None, 0, C.byref(n)))
        recs = (BlockRecord * max(1, n.value))()
        if n.value:
            check(lib().mole_read_blocks(self.h, sys, first, recs, n.value, C.byref(n)))
        dt = np.dtype([("block", np.uint64), ("n_walkers", np.uint64), ("mean", np.float64),
                       ("sum_block_mean", np.float64), ("sum_block_mean2", np.float64)])
        return np.frombuffer(bytes(recs), dtype=dt, count=n.value).copy()

    # ---- diffusion Monte Carlo (DmcRunner, src/dmc/src/dmc.rs) ----

    def dmc_init(self, reference_energy, brancher=BRANCH_SR):
        """mole_dmc_init: one trial energy per system (a scalar for all)."""
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(reference_energy, dtype=np.float64),
                                                 (len(self.systems),)))
        check(lib().mole_dmc_init(self.h, _dp(e), brancher))

    def dmc_diffuse(self, time_step, num_iterations, block_size, num_eq_blocks):
        """mole_dmc_diffuse; returns (energies, errors) of system 0 as
        DmcRunner::diffuse does (see dmc_read for the others)."""
        check(lib().mole_dmc_diffuse(self.h, time_step, num_iterations, block_size, num_eq_blocks))
        e, err, _ = self.dmc_read(0)
        return e, err

    def dmc_read(self, sys=0):
        """(energies, errors, reference_energy) of the last diffuse call."""
        n = C.c_size_t()
        eref = C.c_double()
        check(lib().mole_dmc_read(self.h, sys, None, None, 0, C.byref(n), C.byref(eref)))
        e, err = np.zeros(n.value), np.zeros(n.value)
        if n.value:
            check(lib().mole_dmc_read(self.h, sys, _dp(e), _dp(err), n.value, C.byref(n), C.byref(eref)))
        return e, err, eref.value

    def dmc_read_steps(self, sys=0):
        n = C.c_size_t()
        check(lib().mole_dmc_read_steps(self.h, sys, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        if n.value:
            check(lib().mole_dmc_read_steps(self.h, sys, _dp(out), n.value, C.byref(n)))
        return out

    def dmc_read_walkers(self, sys=0):
        W = self.systems[sys].n_walkers
        w, e = np.zeros(W), np.zeros(W)
        check(lib().mole_dmc_read_walkers(self.h, sys, _dp(w), _dp(e), W))
        return w, e

    def set_launch_config(self, variant=VARIANT_AUTO, block_threads=0):
        check(lib().mole_set_launch_config(self.h, variant, block_threads))

    def set_walker_lanes(self, lanes, sys=0):
        check(lib().mole_set_walker_lanes(self.h, sys, lanes))

    def walker_lanes(self, sys=0):
        v = C.c_uint32()
        check(lib().mole_get_walker_lanes(self.h, sys, C.byref(v)))
        return v.value

    def launch_info(self):
        wg, lds = C.c_uint32(), C.c_size_t()
        check(lib().mole_get_launch_info(self.h, C.byref(wg), C.byref(lds)))
        return wg.value, lds.value


    def set_tile_schedule(self, schedule, workgroups=0):
        """mole_set_tile_schedule: SCHEDULE_CLASSES or SCHEDULE_PERSISTENT."""
        check(lib().mole_set_tile_schedule(self.h, schedule, workgroups))
    def capture(self, n_sweeps, sample=False):
        """HIP graph of one launch of n_sweeps sweeps (see mole_graph_capture)."""
        g = C.c_void_p()
        check(lib().mole_graph_capture(self.h, n_sweeps, SWEEP_SAMPLE if sample else 0, C.byref(g)))
        return Graph(self, g)

    def synchronize(self):
        check(lib().mole_synchronize(self.h))

    def last_timing(self):
        ms, n = C.c_double(), C.c_uint32()
        check(lib().mole_last_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Graph:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def launch(self, replays=1):
        check(lib().mole_graph_launch(self.ctx.h, self.h, replays))

    def close(self):
        if self.h:
            lib().mole_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info():
    buf = C.create_string_buffer(256)
    cu = C.c_int32()
    check(lib().mole_device_info(buf, 256, C.byref(cu)))
    return buf.value.decode(), cu.value


def debug_philox(seed, sys, walker, epoch, first, count):
    out = np.zeros(count, dtype=np.uint64)
    check(lib().mole_debug_philox(seed, sys, walker, epoch, first, count, _u64p(out)))
    return out


def debug_math(op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
    check(lib().mole_debug_math(op, a.size, _dp(a), None if bb is None else _dp(bb), _dp(out)))
    return out


class Mlp:
    """The learned proposal's network (device weights, Adam state, training set)."""

    def __init__(self, seed=2):
        self.h = _P()
        check(lib().mole_mlp_create(seed, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().mole_mlp_destroy(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def weights(self):
        w = np.zeros(MLP_PARAMS, np.float32)
        check(lib().mole_mlp_get_weights(self.h, _fp(w), w.size))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(lib().mole_mlp_set_weights(self.h, _fp(w), w.size))

    def attach(self, ctx):
        check(lib().mole_mlp_attach(ctx.h, self.h))

    def collect(self, ctx, n_snapshots, sweeps_between, sys=0):
        check(lib().mole_mlp_collect(self.h, ctx.h, sys, n_snapshots, sweeps_between))

    def clear_data(self):
        check(lib().mole_mlp_clear_data(self.h))

    def data_size(self):
        n = C.c_size_t()
        check(lib().mole_mlp_data_size(self.h, C.byref(n)))
        return n.value

    def train(self, ctx, steps, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=11, sys=0,
              objective=MLP_OBJ_ESJD):
        d = TrainDesc(steps=steps, batch=batch, lr=lr, beta1=beta1, beta2=beta2, eps=eps, seed=seed,
                      objective=objective)
        obj = np.zeros(max(1, steps))
        check(lib().mole_mlp_train(self.h, ctx.h, sys, C.byref(d), _dp(obj)))
        return obj[:steps]

    def gradient(self, ctx, batch, first=0, seed=11, sys=0, objective=MLP_OBJ_ESJD):
        g = np.zeros(MLP_PARAMS, np.float32)
        o = C.c_double()
        check(lib().mole_mlp_gradient(self.h, ctx.h, sys, batch, first, seed, objective, _fp(g), C.byref(o)))
        return g, o.value

    def eval(self, cfg, nuclei, electron):
        cfg = np.ascontiguousarray(cfg, dtype=np.float64)
        n, threeN = cfg.shape
        nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 3)
        el = np.ascontiguousarray(electron, dtype=np.int32)
        out = np.zeros((n, 4), np.float32)
        check(lib().mole_mlp_eval(self.h, _dp(cfg), threeN // 3, _dp(nuc), nuc.shape[0],
                                  el.ctypes.data_as(C.POINTER(C.c_int32)), n, _fp(out)))
        return out


def debug_mfma_bf16(A, B):
    """One 16x16x16 bf16 MFMA of FP32 16x16 operands (rounded to bf16 on the device)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    assert A.shape == (16, 16) and B.shape == (16, 16)
    D = np.zeros((16, 16), np.float32)
    check(lib().mole_debug_mfma_bf16(_fp(A), _fp(B), _fp(D)))
    return D


def debug_mfma_bf16_k32(A, B):
    """One 16x16x32 bf16 MFMA: A (16x32) times B (32x16), FP32 in, rounded to bf16."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    assert A.shape == (16, 32) and B.shape == (32, 16)
    D = np.zeros((16, 16), np.float32)
    check(lib().mole_debug_mfma_bf16_k32(_fp(A), _fp(B), _fp(D)))
    return D


def debug_mlp_features(cfg, nuclei, electron):
    """mole_debug_mlp_features: the learned proposal's 13 network inputs of
    each (configuration [N][3], electron) sample, as float32 (n, 16)."""
    cfg = np.ascontiguousarray(cfg, dtype=np.float64)
    n = cfg.shape[0]
    N = cfg.reshape(n, -1).shape[1] // 3
    nuc = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 3)
    el = np.ascontiguousarray(electron, dtype=np.int32)
    out = np.zeros((n, 16), np.float32)
    check(lib().mole_debug_mlp_features(_dp(cfg), N, _dp(nuc), nuc.shape[0],
                                        el.ctypes.data_as(C.POINTER(C.c_int32)), n, _fp(out)))
    return out


def parameter_update(sums, optimizer=OPT_SR, step=0.05, sr_mode=SR_TEXTBOOK, sr_eps=1e-3):
    """mole_compute_parameter_update on an SrSums; returns (dp, g)."""
    n = sums.n_params
    dp = np.zeros(n)
    g = np.zeros(n)
    check(lib().mole_compute_parameter_update(C.byref(sums), optimizer, step, sr_mode, sr_eps,
                                              _dp(dp), _dp(g)))
    return dp, g


class Optimizer:
    """A stateful optimizer (mole_optimizer_init / mole_optimizer_update):
    SteepestDescent, StochasticReconfiguration, MomentumDescent,
    NesterovMomentum or OnlineLbfgs of src/optimize/src/optimizers.rs."""

    def __init__(self, kind, n_params, step, momentum=0.0, history=0, sr_mode=SR_TEXTBOOK, sr_eps=1e-3):
        self.state = OptimizerState()
        check(lib().mole_optimizer_init(C.byref(self.state), kind, n_params, step, momentum, history, sr_mode,
                                        sr_eps))

    def compute_parameter_update(self, pars, sums):
        """Returns (dp, g) and advances the state."""
        n = self.state.n_params
        dp, g = np.zeros(n), np.zeros(n)
        pa = np.ascontiguousarray(pars, dtype=np.float64)
        check(lib().mole_optimizer_update(C.byref(self.state), C.byref(sums), _dp(pa), _dp(dp), _dp(g)))
        return dp, g


def debug_copy(n_doubles, repeats=1):
    """Device copy of n_doubles doubles (calibration of HBM byte counters); ms per copy."""
    ms = C.c_double()
    check(lib().mole_debug_copy(n_doubles, repeats, C.byref(ms)))
    return ms.value
