"""ctypes binding of the CPU oracle (liboracle.so) -- test infrastructure only.

Only tests/, the smoke check and the cpu_baseline leg of bench.py use this
module, as the checker.  The product path never imports it.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

NCONS = 5
NGROW = 4


class OParm(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "eos_gamma", "eos_m", "Patm", "cv", "cp", "Ru", "Rg", "kOverMu", "Pr", "C_s", "T_s",
        "smallr", "smallp", "eps", "twoThirds", "oneSix")]


class OFab(ctypes.Structure):
    _fields_ = [("p", ctypes.POINTER(ctypes.c_double)), ("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3),
                ("ncomp", ctypes.c_int)]


class OGeom(ctypes.Structure):
    _fields_ = [
        ("dom_lo", ctypes.c_int * 3), ("dom_hi", ctypes.c_int * 3),
        ("prob_lo", ctypes.c_double * 3), ("dx", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3), ("lo_bc", ctypes.c_int * 3), ("hi_bc", ctypes.c_int * 3),
        ("probtype", ctypes.c_int),
        ("left_state", ctypes.c_double * 5), ("right_state", ctypes.c_double * 5),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_parm_default.argtypes = [ctypes.POINTER(OParm)]
        L.orc_parm_initialize.argtypes = [ctypes.POINTER(OParm)]
        L.orc_c2prim_many.argtypes = [ctypes.c_long, dp, dp, ctypes.POINTER(OParm)]
        L.orc_flux_split_many.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, dp, ctypes.POINTER(OParm)]
        L.orc_weno_many.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp]
        L.orc_compute_dsdt_box.argtypes = [ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OFab), dp, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(OParm), dp, dp, dp]
        L.orc_estdt.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), dp, ctypes.c_double,
                                ctypes.POINTER(OParm)]
        L.orc_estdt.restype = ctypes.c_double
        L.orc_fill_ghost.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OGeom),
                                     ctypes.c_double]
        L.orc_flow_advance.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OFab),
                                       ctypes.POINTER(OFab), ctypes.POINTER(OGeom), ctypes.POINTER(OParm),
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_run.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OFab), ctypes.POINTER(OFab),
                              ctypes.POINTER(OGeom), ctypes.POINTER(OParm), ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        for name in ("orc_init_sod", "orc_init_dmr"):
            getattr(L, name).argtypes = [ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OGeom), ctypes.POINTER(OParm)]
        L.orc_init_kh3d.argtypes = [ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OGeom), ctypes.POINTER(OParm),
                                    ctypes.c_uint64]
        L.orc_dmr_states.argtypes = [dp, dp]
        L.orc_rng_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_rng_uniform.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def parm(**kw):
    p = OParm()
    lib().orc_parm_default(ctypes.byref(p))
    if kw:
        for k, v in kw.items():
            setattr(p, k, v)
        lib().orc_parm_initialize(ctypes.byref(p))
    return p


def geom(n_cell, prob_lo, prob_hi, periodic=(0, 0, 0), lo_bc=(2, 2, 2), hi_bc=(2, 2, 2), probtype=0,
         left_state=None, right_state=None):
    g = OGeom()
    for d in range(3):
        g.dom_lo[d] = 0
        g.dom_hi[d] = n_cell[d] - 1
        g.prob_lo[d] = prob_lo[d]
        g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d]
        g.periodic[d] = periodic[d]
        g.lo_bc[d] = lo_bc[d]
        g.hi_bc[d] = hi_bc[d]
    g.probtype = probtype
    if probtype == 1 and left_state is None:
        left_state, right_state = dmr_states()
    if left_state is not None:
        for n in range(5):
            g.left_state[n] = left_state[n]
            g.right_state[n] = right_state[n]
    return g


def dmr_states():
    l = np.zeros(5)
    r = np.zeros(5)
    lib().orc_dmr_states(_dp(l), _dp(r))
    return l, r


class Fab:
    """A numpy-backed FAB (ncomp, nz, ny, nx) over the inclusive box [lo, hi]."""

    def __init__(self, lo, hi, ncomp=NCONS, data=None):
        self.lo = tuple(lo)
        self.hi = tuple(hi)
        shape = (ncomp,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
        self.a = np.zeros(shape) if data is None else np.ascontiguousarray(data, dtype=np.float64).reshape(shape)
        self.c = OFab()
        self.c.p = _dp(self.a)
        for d in range(3):
            self.c.lo[d] = self.lo[d]
            self.c.hi[d] = self.hi[d]
        self.c.ncomp = ncomp

    def valid(self, lo, hi):
        sl = tuple(slice(lo[d] - self.lo[d], hi[d] - self.lo[d] + 1) for d in (2, 1, 0))
        return self.a[(slice(None),) + sl]


def grown(lo, hi, g=NGROW):
    return tuple(v - g for v in lo), tuple(v + g for v in hi)


def _boxarrays(boxes):
    n = len(boxes)
    lo = (ctypes.c_int * (3 * n))(*[v for b in boxes for v in b[0]])
    hi = (ctypes.c_int * (3 * n))(*[v for b in boxes for v in b[1]])
    return n, lo, hi


def _fabarray(fabs):
    arr = (OFab * len(fabs))()
    for m, f in enumerate(fabs):
        arr[m] = f.c
    return arr


def compute_dsdt_box(lo, hi, S, p, dxinv, do_visc=0, do_gravity=0, faces=False):
    """dSdt on box [lo,hi] from FAB S (covering grow(bx,4)); returns (dSdt, faces)."""
    D = Fab(lo, hi)
    lo_c = (ctypes.c_int * 3)(*lo)
    hi_c = (ctypes.c_int * 3)(*hi)
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    fx = fy = fz = None
    fptr = [None, None, None]
    if faces:
        out = []
        for d in range(3):
            h = list(hi)
            h[d] += 1
            out.append(np.zeros((5,) + tuple(h[e] - lo[e] + 1 for e in (2, 1, 0))))
        fx, fy, fz = out
        fptr = [_dp(fx), _dp(fy), _dp(fz)]
    lib().orc_compute_dsdt_box(lo_c, hi_c, ctypes.byref(S.c), ctypes.byref(D.c), _dp(dxi), do_visc, do_gravity,
                               ctypes.byref(p), *fptr)
    return D.a, (fx, fy, fz)


def estdt(boxes, fabs, dx, cfl, p):
    n, lo, hi = _boxarrays(boxes)
    return lib().orc_estdt(n, lo, hi, _fabarray(fabs), _dp(np.asarray(dx, dtype=np.float64)), cfl, ctypes.byref(p))


def fill_ghost(boxes, fabs, g, time):
    n, lo, hi = _boxarrays(boxes)
    return lib().orc_fill_ghost(n, lo, hi, _fabarray(fabs), ctypes.byref(g), time)


def flow_advance(boxes, A, B, C, g, p, time, dt, scheme, do_visc=0, do_gravity=0):
    n, lo, hi = _boxarrays(boxes)
    return lib().orc_flow_advance(n, lo, hi, _fabarray(A), _fabarray(B), _fabarray(C), ctypes.byref(g),
                                  ctypes.byref(p), time, dt, scheme, do_visc, do_gravity)


def run(boxes, A, B, C, g, p, cfl, scheme=2, do_visc=0, do_gravity=0, stop_time=-1.0, nsteps=-1, time=0.0, dt=0.0):
    n, lo, hi = _boxarrays(boxes)
    t = ctypes.c_double(time)
    d = ctypes.c_double(dt)
    steps = lib().orc_run(n, lo, hi, _fabarray(A), _fabarray(B), _fabarray(C), ctypes.byref(g), ctypes.byref(p), cfl,
                          scheme, do_visc, do_gravity, stop_time, nsteps, ctypes.byref(t), ctypes.byref(d))
    return steps, t.value, d.value


def init_case(kind, boxes, fabs, g, p, seed=0):
    for (lo, hi), f in zip(boxes, fabs):
        lo_c = (ctypes.c_int * 3)(*lo)
        hi_c = (ctypes.c_int * 3)(*hi)
        if kind == 0:
            lib().orc_init_sod(lo_c, hi_c, ctypes.byref(f.c), ctypes.byref(g), ctypes.byref(p))
        elif kind == 1:
            lib().orc_init_dmr(lo_c, hi_c, ctypes.byref(f.c), ctypes.byref(g), ctypes.byref(p))
        else:
            lib().orc_init_kh3d(lo_c, hi_c, ctypes.byref(f.c), ctypes.byref(g), ctypes.byref(p), seed)


def new_level_fabs(boxes, g=NGROW):
    return [Fab(*grown(lo, hi, g)) for lo, hi in boxes]


# ---------------------------------------------------------------------------
# embedded boundary
# ---------------------------------------------------------------------------
class OEbFab(ctypes.Structure):
    _fields_ = [("flag", ctypes.POINTER(ctypes.c_uint32)), ("vfrac", ctypes.POINTER(ctypes.c_double)),
                ("apx", ctypes.POINTER(ctypes.c_double)), ("apy", ctypes.POINTER(ctypes.c_double)),
                ("apz", ctypes.POINTER(ctypes.c_double)), ("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3)]


class EbFab:
    """EB fields of one box over [lo, hi] (numpy arrays (nz, ny, nx)): flag
    (uint32 EBCellFlag words), vfrac, apx, apy, apz."""

    def __init__(self, lo, hi, flag, vfrac, apx, apy, apz):
        self.lo, self.hi = tuple(lo), tuple(hi)
        shape = tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
        self.flag = np.ascontiguousarray(flag, dtype=np.uint32).reshape(shape)
        self.arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(shape) for a in (vfrac, apx, apy, apz)]
        self.c = OEbFab()
        self.c.flag = self.flag.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.c.vfrac, self.c.apx, self.c.apy, self.c.apz = (_dp(a) for a in self.arrs)
        for d in range(3):
            self.c.lo[d] = self.lo[d]
            self.c.hi[d] = self.hi[d]


def _eb_sigs():
    L = lib()
    if getattr(L, "_eb_ready", False):
        return L
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_compute_dsdt_box_eb.argtypes = [ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OFab), dp,
                                          ctypes.POINTER(OEbFab), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(OParm)]
    L.orc_eb_box_kind.argtypes = [ip, ip, ctypes.POINTER(OEbFab)]
    L.orc_eb_set.argtypes = [ctypes.c_int, ctypes.POINTER(OEbFab), ctypes.c_int]
    L.orc_init_eb_inflow.argtypes = [ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(OParm)]
    L.orc_inflow_state.argtypes = [dp]
    L.orc_eb_flux_many.argtypes = [ctypes.c_int, ctypes.c_long, dp, dp, dp, ctypes.POINTER(OParm)]
    L.orc_eb_recon_box.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OEbFab), dp, dp]
    L.orc_eb_visc_box.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(OFab), ctypes.POINTER(OEbFab), dp,
                                  ctypes.POINTER(OParm), dp]
    L.orc_eb_wallflux_many.argtypes = [ctypes.c_long, dp, dp, dp, ctypes.POINTER(OParm)]
    L._eb_ready = True
    return L


def compute_dsdt_box_eb(lo, hi, S, p, dxinv, eb, do_visc=0, do_gravity=0, do_redist=0):
    L = _eb_sigs()
    D = Fab(lo, hi)
    lo_c = (ctypes.c_int * 3)(*lo)
    hi_c = (ctypes.c_int * 3)(*hi)
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    L.orc_compute_dsdt_box_eb(lo_c, hi_c, ctypes.byref(S.c), ctypes.byref(D.c), _dp(dxi), ctypes.byref(eb.c),
                              do_visc, do_gravity, do_redist, ctypes.byref(p))
    return D.a


def eb_flux(d, ql, qr, p):
    """Van Leer flux of (left, right) primitive states (n x 7 each)."""
    L = _eb_sigs()
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    F = np.zeros((ql.shape[0], 5))
    L.orc_eb_flux_many(d, ql.shape[0], _dp(ql), _dp(qr), _dp(F), ctypes.byref(p))
    return F


def eb_recon_box(d, lo, hi, q, eb):
    """Left/right face states of eb_recon on surroundingNodes([lo,hi], d) for
    every component of FAB q: (ql, qr), each (ncomp, faces)."""
    L = _eb_sigs()
    nc = q.a.shape[0]
    ql = np.zeros(_face_shape(lo, hi, d, nc))
    qr = np.zeros_like(ql)
    L.orc_eb_recon_box(d, (ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(q.c), ctypes.byref(eb.c),
                       _dp(ql), _dp(qr))
    return ql, qr


def eb_visc_box(d, lo, hi, q, eb, dxinv, p):
    """Viscous/heat part of the EB face flux (eb_compute_visc) on
    surroundingNodes([lo,hi], d) from the 7-component primitive FAB q."""
    L = _eb_sigs()
    F = np.zeros(_face_shape(lo, hi, d, 5))
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    L.orc_eb_visc_box(d, (ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(q.c), ctypes.byref(eb.c),
                      _dp(dxi), ctypes.byref(p), _dp(F))
    return F


def eb_wallflux(qw, area, p):
    """eb_wallflux of n cells: qw (n x 7 primitives), area (n x 6: apx lo/hi,
    apy lo/hi, apz lo/hi) -> (n x 5)."""
    L = _eb_sigs()
    qw = np.ascontiguousarray(qw, dtype=np.float64)
    area = np.ascontiguousarray(area, dtype=np.float64)
    out = np.zeros((qw.shape[0], 5))
    L.orc_eb_wallflux_many(qw.shape[0], _dp(qw), _dp(area), _dp(out), ctypes.byref(p))
    return out


def eb_box_kind(lo, hi, eb):
    L = _eb_sigs()
    return L.orc_eb_box_kind((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(eb.c))


_eb_keep = None


def set_eb(ebfabs, do_redist=0):
    """EB fields per box for flow_advance / run (None or [] clears them)."""
    global _eb_keep
    L = _eb_sigs()
    if not ebfabs:
        _eb_keep = None
        L.orc_eb_set(0, None, 0)
        return
    arr = (OEbFab * len(ebfabs))()
    for m, e in enumerate(ebfabs):
        arr[m] = e.c
    _eb_keep = (arr, list(ebfabs))
    L.orc_eb_set(len(ebfabs), arr, do_redist)


def init_eb_inflow(boxes, fabs, valid_flags, p):
    L = _eb_sigs()
    for (lo, hi), f, fl in zip(boxes, fabs, valid_flags):
        fl = np.ascontiguousarray(fl, dtype=np.uint32)
        L.orc_init_eb_inflow((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(f.c),
                             fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(p))


def inflow_state():
    L = _eb_sigs()
    st = np.zeros(5)
    L.orc_inflow_state(_dp(st))
    return st


# ---------------------------------------------------------------------------
# chemistry (Li-Dryer H2/air, 9 species: H2 O2 H2O H O OH HO2 H2O2 N2)
# ---------------------------------------------------------------------------
NSPEC = 9
SPECIES = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]


def _chem_sigs():
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    if getattr(L, "_chem_ready", False):
        return L
    L.orc_chem_tables.argtypes = [dp, dp, dp]
    L.orc_production_rate.argtypes = [dp, ctypes.c_double, dp, dp]
    L.orc_gauss.argtypes = [dp, dp, dp]
    L.orc_T_given_ey.argtypes = [ctypes.c_double, dp, ip]
    L.orc_T_given_ey.restype = ctypes.c_double
    for nm in ("orc_CKUBMS", "orc_CKCVBS", "orc_CKCPBS", "orc_CKGAMMA"):
        getattr(L, nm).argtypes = [ctypes.c_double, dp]
        getattr(L, nm).restype = ctypes.c_double
    L.orc_CKPY.argtypes = [dp, ctypes.c_double]
    L.orc_CKPY.restype = ctypes.c_double
    L.orc_CKRHOY.argtypes = [ctypes.c_double, ctypes.c_double, dp]
    L.orc_CKRHOY.restype = ctypes.c_double
    L.orc_CKXTY.argtypes = [dp, dp]
    L.orc_CKYTX.argtypes = [dp, dp]
    L.orc_species_enthalpy.argtypes = [ctypes.c_double, dp]
    L.orc_mixture_properties.argtypes = [ctypes.c_double, dp, dp, dp]
    L.orc_mix_diff_coeffs_mass.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
    L.orc_chem_step_many.argtypes = [ctypes.c_long, dp, dp, ctypes.c_double, ctypes.c_double]
    L.orc_renorm_many.argtypes = [ctypes.c_long, dp, dp]
    L.orc_reactor_init.argtypes = [ctypes.c_long, dp, dp, dp, dp, dp]
    L.orc_reactor_fixed.argtypes = [ctypes.c_long, dp, dp, ctypes.c_double, ctypes.c_int]
    L.orc_reactor_adaptive.argtypes = [ctypes.c_long, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ip, ip]
    L.orc_reactor_T.argtypes = [ctypes.c_long, dp, dp, dp]
    L.orc_reactor_rosenbrock.argtypes = [ctypes.c_long, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ip, ip]
    L.orc_rodas3_fixed.argtypes = [dp, dp, ctypes.c_double, ctypes.c_int]
    L.orc_reactor_rhs.argtypes = [ctypes.c_double, dp, dp, dp, dp]
    L.orc_jacobian_fixed_T.argtypes = [dp, ctypes.c_double, dp]
    L._chem_ready = True
    return L


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def chem_tables():
    mw = np.zeros(9); lo = np.zeros(63); hi = np.zeros(63)
    _chem_sigs().orc_chem_tables(_dp(mw), _dp(lo), _dp(hi))
    return mw, lo.reshape(9, 7), hi.reshape(9, 7)


def production_rate(c, T):
    c = _c(c)
    w = np.zeros(9); J = np.zeros(81)
    _chem_sigs().orc_production_rate(_dp(c), T, _dp(w), _dp(J))
    return w, J.reshape(9, 9)


def gauss(A, b):
    A = _c(A); b = _c(b); x = np.zeros(9)
    _chem_sigs().orc_gauss(_dp(A), _dp(b), _dp(x))
    return x


def T_given_ey(e, rhoi):
    it = ctypes.c_int()
    T = _chem_sigs().orc_T_given_ey(e, _dp(_c(rhoi)), ctypes.byref(it))
    return T, it.value


def thermo(name, *args):
    L = _chem_sigs()
    conv = [_dp(_c(a)) if isinstance(a, (list, np.ndarray)) else a for a in args]
    return getattr(L, "orc_" + name)(*conv)


def CKXTY(x):
    y = np.zeros(9)
    _chem_sigs().orc_CKXTY(_dp(_c(x)), _dp(y))
    return y


def CKYTX(y):
    x = np.zeros(9)
    _chem_sigs().orc_CKYTX(_dp(_c(y)), _dp(x))
    return x


def chem_step(u, rhoi, dt, smallr=1e-19):
    u = _c(u).copy(); r = _c(rhoi).copy()
    _chem_sigs().orc_chem_step_many(len(u) if u.ndim == 2 else 1, _dp(u), _dp(r), dt, smallr)
    return u, r


def renorm(u, rhoi):
    u = _c(u); r = _c(rhoi).copy()
    _chem_sigs().orc_renorm_many(len(u) if u.ndim == 2 else 1, _dp(u), _dp(r))
    return r


def reactor_init(T0, p0, phi):
    T0, p0, phi = _c(T0), _c(p0), _c(phi)
    n = len(T0)
    u = np.zeros((n, 5)); r = np.zeros((n, 9))
    _chem_sigs().orc_reactor_init(n, _dp(T0), _dp(p0), _dp(phi), _dp(u), _dp(r))
    return u, r


def reactor_fixed(u, r, dt, nsteps):
    u = _c(u).copy(); r = _c(r).copy()
    _chem_sigs().orc_reactor_fixed(len(u), _dp(u), _dp(r), dt, nsteps)
    return u, r


def reactor_adaptive(u, r, t_end, h0=1e-8, rtol=1e-6, atol=1e-12, max_steps=200000):
    u = _c(u).copy(); r = _c(r).copy()
    n = len(u)
    steps = np.zeros(n, dtype=np.int32); rej = np.zeros(n, dtype=np.int32)
    _chem_sigs().orc_reactor_adaptive(n, _dp(u), _dp(r), t_end, h0, rtol, atol, max_steps,
                                      steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      rej.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return u, r, steps, rej


def reactor_rosenbrock(u, r, t_end, h0=1e-8, rtol=1e-4, atol=1e-9, max_steps=200000):
    """RODAS3 integration of each reactor to t_end (u, r copied); returns
    (u, r, accepted steps, rejected steps)."""
    u = _c(u).copy(); r = _c(r).copy()
    n = len(u)
    steps = np.zeros(n, dtype=np.int32); rej = np.zeros(n, dtype=np.int32)
    _chem_sigs().orc_reactor_rosenbrock(n, _dp(u), _dp(r), t_end, h0, rtol, atol, max_steps,
                                        steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        rej.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return u, r, steps, rej


def rodas3_fixed(u1, r1, h, nsteps):
    """nsteps RODAS3 steps of h of one reactor (u1: 5, r1: 9), no error control."""
    u = _c(u1).copy(); r = _c(r1).copy()
    _chem_sigs().orc_rodas3_fixed(_dp(u), _dp(r), h, nsteps)
    return u, r


def reactor_rhs(E0, y0, y):
    """d rho/dt of the reactor ODE at densities y (energy from E0 at y0) and the temperature there."""
    y0 = _c(y0); y = _c(y); f = np.zeros(9); T = np.zeros(1)
    _chem_sigs().orc_reactor_rhs(E0, _dp(y0), _dp(y), _dp(f), _dp(T))
    return f, T[0]


def jacobian_fixed_T(c, T):
    """The complete d wdot/d c (CGS) at fixed T."""
    c = _c(c); J = np.zeros(81)
    _chem_sigs().orc_jacobian_fixed_T(_dp(c), T, _dp(J))
    return J.reshape(9, 9)


def reactor_T(u, r):
    u = _c(u); r = _c(r); T = np.zeros(len(u))
    _chem_sigs().orc_reactor_T(len(u), _dp(u), _dp(r), _dp(T))
    return T


# the reference's own AMReX-free kinetics (Test/chemical_test.cpp), built by oracle/ref.mk
REF_DIR = HERE / "_ref"
_ref = None


_ref_t = None


def ref_trans():
    """ctypes handle of oracle/_ref/libtransref.so (Test/transport_test.cpp), or None."""
    global _ref_t
    if _ref_t is None:
        p = REF_DIR / "libtransref.so"
        if not p.exists():
            return None
        L = ctypes.CDLL(str(p))
        dp = ctypes.POINTER(ctypes.c_double)
        L.mixtureProperties = L._Z17mixturePropertiesdPdRdS0_
        L.mixtureProperties.argtypes = [ctypes.c_double, dp, dp, dp]
        L.getMixDiffCoeffsMass = L._Z20getMixDiffCoeffsMassddPdS_
        L.getMixDiffCoeffsMass.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        _ref_t = L
    return _ref_t


def mixture_properties(T, X):
    L = _chem_sigs()
    X = np.ascontiguousarray(X, dtype=np.float64)
    mu = ctypes.c_double(); lam = ctypes.c_double()
    L.orc_mixture_properties(T, _dp(X), ctypes.byref(mu), ctypes.byref(lam))
    return mu.value, lam.value


def mix_diff_coeffs(T, p, X):
    L = _chem_sigs()
    X = np.ascontiguousarray(X, dtype=np.float64)
    D = np.zeros(NSPEC)
    L.orc_mix_diff_coeffs_mass(T, p, _dp(X), _dp(D))
    return D


def ref_chem():
    """ctypes handle of oracle/_ref/libchemref.so, or None when it was not built."""
    global _ref
    if _ref is None:
        p = REF_DIR / "libchemref.so"
        if not p.exists():
            return None
        L = ctypes.CDLL(str(p))
        dp = ctypes.POINTER(ctypes.c_double)
        L.vproductionRate = L._Z15vproductionRatePdRA9_A9_dS_d
        L.vproductionRate.argtypes = [dp, dp, dp, ctypes.c_double]
        L.GET_T_GIVEN_EY = L._Z14GET_T_GIVEN_EYdPdRd
        L.GET_T_GIVEN_EY.argtypes = [ctypes.c_double, dp, dp]
        L.gauss = L._Z5gaussPdRA9_A9_dS_
        L.gauss.argtypes = [dp, dp, dp]
        for nm, mang in (("CKUBMS", "_Z6CKUBMSdPdRd"), ("CKCVBS", "_Z6CKCVBSdPdRd"), ("CKCPBS", "_Z6CKCPBSdPdRd"),
                         ("CKGAMMA", "_Z7CKGAMMAdPdRd")):
            f = getattr(L, mang)
            f.argtypes = [ctypes.c_double, dp, dp]
            setattr(L, nm, f)
        L.CKPY = L._Z4CKPYPddRd
        L.CKPY.argtypes = [dp, ctypes.c_double, dp]
        L.CKRHOY = L._Z6CKRHOYddPdRd
        L.CKRHOY.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        L.CKXTY = L._Z5CKXTYPdS_
        L.CKXTY.argtypes = [dp, dp]
        L.CKYTX = L._Z5CKYTXPdS_
        L.CKYTX.argtypes = [dp, dp]
        _ref = L
    return _ref


# ---------------------------------------------------------------------------
# multispecies (CHEM) flow
# ---------------------------------------------------------------------------
def _multi_sigs():
    L = lib()
    if getattr(L, "_multi_ready", False):
        return L
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    F = ctypes.POINTER(OFab)
    L.orc_compute_dsdt_multi_box.argtypes = [ip, ip, F, F, F, F, dp, ctypes.c_int, ctypes.POINTER(OParm)]
    L.orc_fill_ghost_spec.argtypes = [ctypes.c_int, ip, ip, F, ctypes.POINTER(OGeom), ctypes.c_double]
    L.orc_flow_advance_multi.argtypes = [ctypes.c_int, ip, ip, F, F, F, F, ctypes.POINTER(OGeom), ctypes.POINTER(OParm),
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.orc_advance_chem.argtypes = [ctypes.c_int, ip, ip, F, F, F, F, ctypes.POINTER(OGeom), ctypes.POINTER(OParm),
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int]
    L.orc_init_chemsod.argtypes = [ip, ip, F, F, ctypes.POINTER(OGeom)]
    L.orc_compute_dsdt_multi_box_eb.argtypes = [ip, ip, F, F, F, F, dp, ctypes.POINTER(OEbFab), ctypes.c_int,
                                                ctypes.POINTER(OParm)]
    L.orc_compute_dsdt_multi_box_faces.argtypes = [ip, ip, F, F, dp, ctypes.c_int, ctypes.POINTER(OParm),
                                                   ctypes.POINTER(dp)]
    L.orc_vl_species_flux_many.argtypes = [ctypes.c_long, dp, dp, dp, dp, dp, dp, dp]
    L.orc_set_inflow_species.argtypes = [dp]
    L._multi_ready = True
    return L


def compute_dsdt_multi_box(lo, hi, S, Sp, p, dxinv, do_visc=0):
    L = _multi_sigs()
    D = Fab(lo, hi)
    Ds = Fab(lo, hi, ncomp=NSPEC)
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    L.orc_compute_dsdt_multi_box((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(S.c),
                                 ctypes.byref(Sp.c), ctypes.byref(D.c), ctypes.byref(Ds.c), _dp(dxi), do_visc,
                                 ctypes.byref(p))
    return D.a, Ds.a


def _face_shape(lo, hi, d, ncomp):
    h = list(hi)
    h[d] += 1
    return (ncomp,) + tuple(h[e] - lo[e] + 1 for e in (2, 1, 0))


def compute_multi_faces(lo, hi, S, Sp, p, dxinv, do_visc=0):
    """Face fluxes of compute_dSdt_multi on one regular box: ([F_x, F_y, F_z]
    (5 comps), [species F_x..z] (9), [Fickian J_x..z] (9)), each over
    surroundingNodes(bx, d)."""
    L = _multi_sigs()
    outs = [[np.zeros(_face_shape(lo, hi, d, nc)) for d in range(3)] for nc in (5, NSPEC, NSPEC)]
    ptrs = (ctypes.POINTER(ctypes.c_double) * 9)(*[_dp(a) for grp in outs for a in grp])
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    L.orc_compute_dsdt_multi_box_faces((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(S.c),
                                       ctypes.byref(Sp.c), _dp(dxi), do_visc, ctypes.byref(p), ptrs)
    return outs


_spec_in_keep = None


def set_inflow_species(spec):
    """Species the ghost fill writes on ext_dir sides (None clears)."""
    global _spec_in_keep
    L = _multi_sigs()
    if spec is None:
        _spec_in_keep = None
        L.orc_set_inflow_species(None)
        return
    _spec_in_keep = np.ascontiguousarray(spec, dtype=np.float64)
    L.orc_set_inflow_species(_dp(_spec_in_keep))


def vl_species_flux(ul, ur, cL, cR, rl, rr):
    """Van Leer mass-flux split of the partial densities on n faces."""
    L = _multi_sigs()
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (ul, ur, cL, cR, rl, rr)]
    n = arr[0].shape[0]
    F = np.zeros((n, NSPEC))
    L.orc_vl_species_flux_many(n, *[_dp(a) for a in arr], _dp(F))
    return F


def compute_dsdt_multi_box_eb(lo, hi, S, Sp, p, dxinv, eb, do_visc=0):
    L = _multi_sigs()
    D = Fab(lo, hi)
    Ds = Fab(lo, hi, ncomp=NSPEC)
    dxi = np.ascontiguousarray(dxinv, dtype=np.float64)
    L.orc_compute_dsdt_multi_box_eb((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(S.c),
                                    ctypes.byref(Sp.c), ctypes.byref(D.c), ctypes.byref(Ds.c), _dp(dxi),
                                    ctypes.byref(eb.c), do_visc, ctypes.byref(p))
    return D.a, Ds.a


def fill_ghost_spec(boxes, fabs, g, time):
    L = _multi_sigs()
    n, lo, hi = _boxarrays(boxes)
    return L.orc_fill_ghost_spec(n, lo, hi, _fabarray(fabs), ctypes.byref(g), time)


def flow_advance_multi(boxes, A, As, B, Bs, g, p, time, dt, scheme=1, do_visc=0):
    L = _multi_sigs()
    n, lo, hi = _boxarrays(boxes)
    return L.orc_flow_advance_multi(n, lo, hi, _fabarray(A), _fabarray(As), _fabarray(B), _fabarray(Bs),
                                    ctypes.byref(g), ctypes.byref(p), time, dt, scheme, do_visc)


def advance_chem(boxes, A, As, B, Bs, g, p, time, dt, level=0, scheme=1, do_visc=0, renorm=1):
    L = _multi_sigs()
    n, lo, hi = _boxarrays(boxes)
    return L.orc_advance_chem(n, lo, hi, _fabarray(A), _fabarray(As), _fabarray(B), _fabarray(Bs),
                              ctypes.byref(g), ctypes.byref(p), time, dt, level, scheme, do_visc, renorm)


def init_chemsod(boxes, fabs, sfabs, g):
    L = _multi_sigs()
    for (lo, hi), f, fs in zip(boxes, fabs, sfabs):
        L.orc_init_chemsod((ctypes.c_int * 3)(*lo), (ctypes.c_int * 3)(*hi), ctypes.byref(f.c), ctypes.byref(fs.c),
                           ctypes.byref(g))


def new_spec_fabs(boxes, g=NGROW):
    return [Fab(*grown(lo, hi, g), ncomp=NSPEC) for lo, hi in boxes]
