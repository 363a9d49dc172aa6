"""ctypes view of the oracle's Slater-Jastrow functions (oracle/sj.h).

Test infrastructure only (tests/, smoke, bench cpu_baseline)."""
import ctypes as C

import numpy as np

from . import oracle_ffi as OF

PREC_FP64 = 0
PREC_MIXED = 1
NPARAM = 6  # five Jastrow parameters and the exponent scale

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


class Sj(C.Structure):
    _fields_ = [
        ("n_up", C.c_int), ("n_dn", C.c_int), ("n_nuc", C.c_int), ("n_basis", C.c_int),
        ("n_orb", C.c_int), ("hamiltonian", C.c_int),
        ("nuc", _pd), ("charge", _pi), ("zeta", _pd), ("center", _pi), ("coef", _pd),
        ("p", C.c_double * NPARAM),
        ("kind", _pi),
        ("opt_exp", C.c_int),
    ]


TERM_KINETIC, TERM_IONIC_POTENTIAL, TERM_ELECTRONIC_POTENTIAL, TERM_HARMONIC, TERM_RADIAL_POWER, TERM_DIPOLE = range(6)


class OrcTerm(C.Structure):
    _fields_ = [("kind", C.c_int), ("coef", C.c_double), ("param", C.c_double * 6)]


class OrcOperator(C.Structure):
    _fields_ = [("n_terms", C.c_int), ("term", OrcTerm * 4)]


def operator(terms):
    """[(kind, coef, params...)] -> OrcOperator."""
    op = OrcOperator(n_terms=len(terms))
    for k, t in enumerate(terms):
        op.term[k].kind = int(t[0])
        op.term[k].coef = float(t[1])
        for q, v in enumerate(t[2:]):
            op.term[k].param[q] = float(v)
    return op


class SjRunParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("sys", C.c_uint32), ("epoch", C.c_uint32), ("box_side", C.c_double),
        ("steps", C.c_uint32), ("block_size", C.c_uint32), ("init_from_stream", C.c_int),
        ("pos0", C.c_uint64), ("move0", C.c_uint64), ("precision", C.c_int),
        ("lu_period", C.c_uint32), ("proposal", C.c_int), ("tau", C.c_double),
        ("mlp_w", C.POINTER(C.c_float)), ("mlp_emulate_bf16", C.c_int), ("mlp_drop_q_ratio", C.c_int),
        ("hamiltonian", C.POINTER(OrcOperator)), ("obs", C.POINTER(OrcOperator)), ("n_obs", C.c_int),
    ]


class SjRunOut(C.Structure):
    _fields_ = [
        ("cfg", _pd), ("accept_mask", C.POINTER(C.c_uint64)), ("e_local", _pd), ("dlog", _pd),
        ("sum_e", C.c_double), ("sum_e2", C.c_double), ("sum_bm", C.c_double), ("sum_bm2", C.c_double),
        ("n_samples", C.c_uint64), ("n_blocks", C.c_uint64), ("n_accepted", C.c_uint64),
        ("n_moves", C.c_uint64), ("obs_sum", _pd), ("obs_sum2", _pd),
        ("ld", _pd), ("max_ld_drift", C.c_double), ("n_singular", C.c_uint64), ("psi", _pd),
    ]


_bound = False


def lib():
    global _bound
    L = OF.lib()
    if not _bound:
        ps = C.POINTER(Sj)
        L.orc_sj_value.restype = C.c_double
        L.orc_sj_value.argtypes = [ps, _pd]
        L.orc_sj_logdet.restype = C.c_double
        L.orc_sj_logdet.argtypes = [ps, _pd, C.c_int, _pd]
        L.orc_sj_det_leibniz.restype = C.c_longdouble
        L.orc_sj_det_leibniz.argtypes = [ps, _pd, C.c_int]
        L.orc_sj_inverse.restype = C.c_int
        L.orc_sj_inverse.argtypes = [ps, _pd, C.c_int, _pd]
        L.orc_sj_inverse_det.restype = C.c_int
        L.orc_sj_inverse_det.argtypes = [ps, _pd, C.c_int, _pd, _pd, _pd]
        L.orc_sj_logdet_value.restype = C.c_double
        L.orc_sj_logdet_value.argtypes = [C.c_double, C.c_double]
        L.orc_sj_logdet_mul.restype = None
        L.orc_sj_logdet_mul.argtypes = [_pd, _pd, C.c_double]
        L.orc_sj_local_energy.restype = C.c_double
        L.orc_sj_local_energy.argtypes = [ps, _pd, _pd, _pd, _pd]
        L.orc_sj_local_energy_full.restype = C.c_double
        L.orc_sj_local_energy_full.argtypes = [ps, _pd, _pd]
        L.orc_sj_drift.argtypes = [ps, _pd, C.c_int, _pd]
        L.orc_sj_operator.restype = C.c_double
        L.orc_sj_operator.argtypes = [ps, _pd, _pd, _pd, C.POINTER(OrcOperator)]
        for f in (L.orc_sj_run_parity, L.orc_sj_run_faithful):
            f.restype = C.c_int
            f.argtypes = [ps, C.POINTER(SjRunParams), C.c_uint32, C.c_uint32, C.c_int,
                          C.POINTER(SjRunOut)]
        _bound = True
    return L


class SjWave:
    """Keeps the arrays of a systems.SJSystem alive behind an orc_sj."""

    def __init__(self, system, hamiltonian=0, params=None):
        """params: the five Jastrow parameters, or six with the exponent
        scale s (the basis exponents become s zeta_m, the device's
        products)."""
        s = system
        kind = getattr(s, "kind", None)
        kind = np.zeros(s.zeta.size, np.intc) if kind is None else np.ascontiguousarray(kind, dtype=np.intc)
        p = s.jastrow if params is None else params
        opt_exp = len(p) == NPARAM
        zeta = np.ascontiguousarray(float(p[5]) * s.zeta if opt_exp else s.zeta, dtype=np.float64)
        self._keep = [s.nuclei, s.charges.astype(np.intc), zeta, s.center.astype(np.intc), s.coef, kind]
        nuc, ch, ze, ce, co, kd = self._keep
        self.st = Sj(n_up=s.n_up, n_dn=s.n_dn, n_nuc=nuc.shape[0], n_basis=ze.size, n_orb=co.shape[0],
                     hamiltonian=hamiltonian, nuc=nuc.ctypes.data_as(_pd), charge=ch.ctypes.data_as(_pi),
                     zeta=ze.ctypes.data_as(_pd), center=ce.ctypes.data_as(_pi),
                     coef=co.ctypes.data_as(_pd), kind=kd.ctypes.data_as(_pi))
        for k in range(NPARAM):
            self.st.p[k] = float(p[k]) if k < len(p) else 1.0
        self.st.opt_exp = int(opt_exp)
        self.n_up, self.n_dn = s.n_up, s.n_dn
        self.N = s.n_up + s.n_dn

    def ref(self):
        return C.byref(self.st)


def value(wv, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_sj_value(wv.ref(), OF.dptr(x))


def inverse(wv, x, spin):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = wv.n_up if spin == 0 else wv.n_dn
    B = np.zeros(max(1, n * n))
    rc = lib().orc_sj_inverse(wv.ref(), OF.dptr(x), spin, OF.dptr(B))
    return B[:n * n].reshape(n, n), rc


def inverse_det(wv, x, spin):
    """(inverse, mant, ex, rc): the device's re-inversion and the determinant
    mant 2^ex from its pivots (mant 0 for a singular block)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = wv.n_up if spin == 0 else wv.n_dn
    B = np.zeros(max(1, n * n))
    m, e = C.c_double(), C.c_double()
    rc = lib().orc_sj_inverse_det(wv.ref(), OF.dptr(x), spin, OF.dptr(B), C.byref(m), C.byref(e))
    return B[:n * n].reshape(n, n), m.value, e.value, rc


def logdet_value(mant, ex):
    return lib().orc_sj_logdet_value(float(mant), float(ex))


def logdet(wv, x, spin):
    """(log|det A_s|, sign) by the oracle's LU (an elimination independent of
    the Gauss-Jordan re-inversion)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    sg = C.c_double()
    v = lib().orc_sj_logdet(wv.ref(), OF.dptr(x), int(spin), C.byref(sg))
    return v, sg.value


def drift(wv, x, i):
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.zeros(3)
    lib().orc_sj_drift(wv.ref(), x.ctypes.data_as(_pd), int(i), v.ctypes.data_as(_pd))
    return v


def local_energy(wv, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dl = np.zeros(NPARAM)
    e = lib().orc_sj_local_energy_full(wv.ref(), OF.dptr(x), OF.dptr(dl))
    return e, dl


BOX, DIFFUSE, LEARNED = 0, 1, 2


def operator_value(wv, x, terms):
    """orc_sj_operator at configuration x (fresh inverses)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    Bu, _ = inverse(wv, x, 0)
    Bd, _ = inverse(wv, x, 1)
    Bu = np.ascontiguousarray(Bu.ravel() if Bu.size else np.zeros(1))
    Bd = np.ascontiguousarray(Bd.ravel() if Bd.size else np.zeros(1))
    op = operator(terms)
    return lib().orc_sj_operator(wv.ref(), OF.dptr(x), OF.dptr(Bu), OF.dptr(Bd), C.byref(op))


def run(wv, *, seed, walkers, steps, block_size, box_side=0.0, faithful=False, precision=PREC_FP64,
        lu_period=0, sys=0, epoch=0, cfg=None, pos0=0, move0=0, n_threads=1, record=True,
        proposal=BOX, tau=0.0, mlp_w=None, emulate_bf16=False, drop_q_ratio=False, hamiltonian=None,
        observables=()):
    L = lib()
    N = wv.N
    rp = SjRunParams(seed=seed, sys=sys, epoch=epoch, box_side=box_side, steps=steps,
                     block_size=block_size, init_from_stream=0 if cfg is not None else 1, pos0=pos0,
                     move0=move0, precision=precision, lu_period=lu_period,
                     proposal=proposal, tau=tau, mlp_emulate_bf16=int(emulate_bf16),
                     mlp_drop_q_ratio=int(drop_q_ratio))
    if mlp_w is not None:
        mlp_w = np.ascontiguousarray(mlp_w, dtype=np.float32)
        rp.mlp_w = mlp_w.ctypes.data_as(C.POINTER(C.c_float))
    keep = []
    if hamiltonian is not None:
        h = operator(hamiltonian)
        keep.append(h)
        rp.hamiltonian = C.pointer(h)
    n_obs = len(observables)
    obs_sum, obs_sum2 = np.zeros(max(1, n_obs)), np.zeros(max(1, n_obs))
    if n_obs:
        arr = (OrcOperator * n_obs)(*[operator(o) for o in observables])
        keep.append(arr)
        rp.obs = C.cast(arr, C.POINTER(OrcOperator))
        rp.n_obs = n_obs
    blocks = steps // block_size
    ns = (blocks - 1) * block_size
    out = SjRunOut()
    out.obs_sum = OF.dptr(obs_sum)
    out.obs_sum2 = OF.dptr(obs_sum2)
    ld = np.zeros((walkers, 4))
    out.ld = OF.dptr(ld)
    c = np.zeros((walkers, 3 * N)) if cfg is None else np.ascontiguousarray(cfg, dtype=np.float64).copy()
    out.cfg = OF.dptr(c)
    res = {}
    if record:
        masks = np.zeros((walkers, blocks * block_size), dtype=np.uint64)
        el = np.zeros((walkers, ns))
        dl = np.zeros((walkers, ns, NPARAM))
        out.accept_mask = OF.u64ptr(masks)
        out.e_local = OF.dptr(el)
        out.dlog = OF.dptr(dl)
        res.update(accept_mask=masks, e_local=el, dlog=dl)
        if not faithful:
            ps = np.zeros((walkers, ns))
            out.psi = OF.dptr(ps)
            res.update(psi=ps)
    f = L.orc_sj_run_faithful if faithful else L.orc_sj_run_parity
    rc = f(wv.ref(), C.byref(rp), 0, walkers, n_threads, C.byref(out))
    if rc != 0:
        raise ValueError("orc_sj_run failed (steps < 2*block_size or sizes too large)")
    res.update(cfg=c, sum_e=out.sum_e, sum_e2=out.sum_e2, sum_bm=out.sum_bm, sum_bm2=out.sum_bm2,
               n_samples=out.n_samples, n_blocks=out.n_blocks, n_accepted=out.n_accepted,
               n_moves=out.n_moves, obs_sum=obs_sum[:n_obs], obs_sum2=obs_sum2[:n_obs],
               ld=ld, max_ld_drift=out.max_ld_drift, n_singular=out.n_singular)
    return res


# ---- diffusion Monte Carlo (orc_dmc_*, orc_branch_*) -----------------------

BRANCH_SR, BRANCH_SIMPLE = 0, 1
DMC_INIT_EPOCH = 0x444D4300
DMC_BRANCH_EPOCH = 0x444D4301


class DmcParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("sys", C.c_uint32), ("epoch", C.c_uint32), ("tau", C.c_double),
        ("precision", C.c_int), ("lu_period", C.c_uint32), ("n_walkers", C.c_uint32), ("brancher", C.c_int),
        ("hamiltonian", C.POINTER(OrcOperator)),
    ]


class DmcState(C.Structure):
    _fields_ = [
        ("cfg", _pd), ("weight", _pd), ("e_local", _pd), ("B", _pd), ("reference_energy", C.c_double),
        ("steps", C.c_uint64), ("draw", C.c_uint64), ("moves", C.c_uint64),
    ]


def _bind_dmc(L):
    if getattr(L, "_dmc_bound", False):
        return
    pu = C.POINTER(C.c_uint32)
    L.orc_branch_sr.restype = C.c_int
    L.orc_branch_sr.argtypes = [_pd, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64, pu, _pd]
    L.orc_branch_simple.restype = None
    L.orc_branch_simple.argtypes = [_pd, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint64, pu, _pd]
    L.orc_dmc_init.restype = C.c_int
    L.orc_dmc_init.argtypes = [C.POINTER(Sj), C.POINTER(DmcParams), C.c_double, C.POINTER(DmcState)]
    L.orc_dmc_diffuse.restype = C.c_int
    L.orc_dmc_diffuse.argtypes = [C.POINTER(Sj), C.POINTER(DmcParams), C.POINTER(DmcState), C.c_double,
                                  C.c_uint32, C.c_uint32, C.c_uint32, _pd, _pd, _pd]
    L._dmc_bound = True


def branch(w, seed, sys, step, brancher):
    """orc_branch_sr / orc_branch_simple: (src indices, new weights)."""
    L = lib()
    _bind_dmc(L)
    w = np.ascontiguousarray(w, dtype=np.float64)
    W = w.size
    src = np.zeros(W, dtype=np.uint32)
    wo = np.zeros(W)
    if brancher == BRANCH_SR:
        rc = L.orc_branch_sr(OF.dptr(w), W, seed, sys, step, src.ctypes.data_as(C.POINTER(C.c_uint32)),
                             OF.dptr(wo))
        if rc != 0:
            raise ValueError("total integer weight 0")
    else:
        L.orc_branch_simple(OF.dptr(w), W, seed, sys, step, src.ctypes.data_as(C.POINTER(C.c_uint32)),
                            OF.dptr(wo))
    return src, wo


class Dmc:
    """The oracle's DmcRunner on one Slater-Jastrow system."""

    def __init__(self, wv, *, walkers, seed, tau, reference_energy, brancher=BRANCH_SR, precision=PREC_FP64,
                 lu_period=0, sys=0, epoch=0, hamiltonian=None):
        L = lib()
        _bind_dmc(L)
        self.wv = wv
        nb = wv.n_up * wv.n_up + wv.n_dn * wv.n_dn
        self.cfg = np.zeros((walkers, 3 * wv.N))
        self.weight = np.zeros(walkers)
        self.e_local = np.zeros(walkers)
        self.B = np.zeros((walkers, max(1, nb)))
        self.p = DmcParams(seed=seed, sys=sys, epoch=epoch, tau=tau, precision=precision, lu_period=lu_period,
                           n_walkers=walkers, brancher=brancher)
        if hamiltonian is not None:
            self._h = operator(hamiltonian)
            self.p.hamiltonian = C.pointer(self._h)
        self.st = DmcState(cfg=OF.dptr(self.cfg), weight=OF.dptr(self.weight), e_local=OF.dptr(self.e_local),
                           B=OF.dptr(self.B))
        assert L.orc_dmc_init(wv.ref(), C.byref(self.p), reference_energy, C.byref(self.st)) == 0

    def diffuse(self, time_step, num_iterations, block_size, num_eq_blocks):
        L = lib()
        nb = max(1, num_iterations // block_size)
        e, err, se = np.zeros(nb), np.zeros(nb), np.zeros(max(1, num_iterations))
        n = L.orc_dmc_diffuse(self.wv.ref(), C.byref(self.p), C.byref(self.st), time_step, num_iterations,
                              block_size, num_eq_blocks, OF.dptr(e), OF.dptr(err), OF.dptr(se))
        if n < 0:
            raise ValueError("orc_dmc_diffuse failed")
        return e[:n], err[:n], se[:num_iterations]

    @property
    def reference_energy(self):
        return self.st.reference_energy
