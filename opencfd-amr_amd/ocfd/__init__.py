"""Python binding of the MI355X-native per-box advance (libocfd.so, C ABI of
include/ocfd.h).

Device memory and streams come from torch (plumbing only); every computation
runs in the HIP library.  Loading fails loudly when the library is missing.
"""
import ctypes
import itertools
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "libocfd.so"

NCONS = 5
NGROW = 4
NSPEC = 9

# status codes of the C ABI
OK, EINVAL, EHIP, ENOMEM, EUNSUPPORTED, ENONFINITE = 0, 1, 2, 3, 4, 5

VISC = 1
GRAVITY = 2
FLUX_FP32 = 4
PLANAR = 8

# physical BC codes (Exec/*/inputs)
INTERIOR, INFLOW, OUTFLOW, SYMMETRY, SLIPWALL, NOSLIPWALL = range(6)

CASE_SOD, CASE_DMR, CASE_KH3D, CASE_EB_INFLOW, CASE_CHEMSOD = 0, 1, 2, 3, 4
CASE_CHEM_CHANNEL = 5   # premixed reacting channel: uniform data from ocfd_init_species_uniform


class Parm(ctypes.Structure):
    """struct Parm, Source/Constants.H:7-64 (non-CHEM members)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "eos_gamma", "eos_m", "Patm", "cv", "cp", "Ru", "Rg", "kOverMu", "Pr", "C_s", "T_s",
        "smallr", "smallp", "eps", "twoThirds", "oneSix")]


class Geom(ctypes.Structure):
    _fields_ = [
        ("dom_lo", ctypes.c_int * 3), ("dom_hi", ctypes.c_int * 3),
        ("prob_lo", ctypes.c_double * 3), ("dx", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3), ("lo_bc", ctypes.c_int * 3), ("hi_bc", ctypes.c_int * 3),
        ("probtype", ctypes.c_int),
        ("left_state", ctypes.c_double * 5), ("right_state", ctypes.c_double * 5),
    ]


EB_NONE, EB_CYLINDER, EB_BOX, EB_PLANE = range(4)
EB_REDIST = 16


class EbSpec(ctypes.Structure):
    """struct ocfd_eb_spec: a z-extruded body for the cut-cell geometry."""
    _fields_ = [("kind", ctypes.c_int), ("fluid_inside", ctypes.c_int), ("center", ctypes.c_double * 3),
                ("radius", ctypes.c_double), ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("point", ctypes.c_double * 3), ("normal", ctypes.c_double * 3), ("small_volfrac", ctypes.c_double)]


def eb_cylinder(cx, cy, r, small_volfrac=0.0):
    s = EbSpec()
    s.kind = EB_CYLINDER
    s.center[0], s.center[1] = cx, cy
    s.radius = r
    s.small_volfrac = small_volfrac
    return s


def eb_box(lo, hi, small_volfrac=0.0):
    s = EbSpec()
    s.kind = EB_BOX
    for d in range(2):
        s.lo[d], s.hi[d] = lo[d], hi[d]
    s.small_volfrac = small_volfrac
    return s


def eb_plane(point, normal, small_volfrac=0.0):
    s = EbSpec()
    s.kind = EB_PLANE
    for d in range(2):
        s.point[d], s.normal[d] = point[d], normal[d]
    s.small_volfrac = small_volfrac
    return s


class StepCtl(ctypes.Structure):
    """Device time control of one level (struct ocfd_step_ctl)."""
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double), ("cfl", ctypes.c_double),
                ("stop_time", ctypes.c_double), ("steps", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("done", ctypes.c_int), ("status", ctypes.c_int), ("pad1", ctypes.c_char * 80),
                ("red", ctypes.c_ulonglong), ("bad", ctypes.c_uint), ("pad2", ctypes.c_char * 116)]


CTL_BYTES = ctypes.sizeof(StepCtl)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP library {LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.ocfd_parm_initialize.argtypes = [ctypes.POINTER(Parm)]
        L.ocfd_level_create.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(Geom), ctypes.POINTER(vp)]
        L.ocfd_level_destroy.argtypes = [vp]
        L.ocfd_level_mf_size.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ocfd_level_mf_size.restype = ctypes.c_long
        L.ocfd_level_fab_offset.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ocfd_level_fab_offset.restype = ctypes.c_long
        L.ocfd_level_nbox.argtypes = [vp]
        L.ocfd_level_create_eb.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(Geom), ctypes.POINTER(EbSpec),
                                           ctypes.POINTER(vp)]
        L.ocfd_level_box_types.argtypes = [vp, ip]
        L.ocfd_eb_geometry.argtypes = [ctypes.POINTER(EbSpec), ctypes.POINTER(Geom), ip, ip, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint32), dp, dp, dp, dp, dp, dp, dp, dp]
        L.ocfd_inflow_state.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.ocfd_level_device_bytes.argtypes = [vp]
        L.ocfd_level_device_bytes.restype = ctypes.c_long
        L.ocfd_cached_device_bytes.argtypes = []
        L.ocfd_cached_device_bytes.restype = ctypes.c_long
        L.ocfd_release_cached_memory.argtypes = []
        L.ocfd_compute_dsdt.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_rk_stage.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_fill_boundary.argtypes = [vp, vp, ctypes.c_double, vp]
        L.ocfd_estdt.argtypes = [vp, vp, ctypes.c_double, ctypes.POINTER(Parm), dp, vp]
        L.ocfd_advance.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_init_state.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(Parm), ctypes.c_ulonglong]
        L.ocfd_dmr_states.argtypes = [dp, dp]
        L.ocfd_profile_stage.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.POINTER(Parm), ctypes.c_int, ctypes.c_int, dp, vp]
        L.ocfd_chem_advance.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.POINTER(Parm), vp]
        L.ocfd_species_renorm.argtypes = [vp, vp, vp, vp]
        L.ocfd_reactor_init.argtypes = [ctypes.c_long, vp, vp, vp, vp, vp]
        L.ocfd_reactor_advance_fixed.argtypes = [ctypes.c_long, vp, ctypes.c_double, ctypes.c_int, vp]
        L.ocfd_reactor_advance_adaptive.argtypes = [ctypes.c_long, vp, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp, vp, vp]
        L.ocfd_reactor_temperature.argtypes = [ctypes.c_long, vp, vp, vp]
        L.ocfd_ctl_start.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(Parm), vp]
        L.ocfd_step_device.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_graph_capture_steps.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, ctypes.POINTER(Parm), ctypes.c_int,
                                               ctypes.c_int, vp, ctypes.POINTER(vp)]
        L.ocfd_graph_launch.argtypes = [vp, vp]
        L.ocfd_graph_destroy.argtypes = [vp]
        L.ocfd_fill_boundary_spec.argtypes = [vp, vp, ctypes.c_double, vp]
        L.ocfd_compute_dsdt_multi.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_compute_dsdt_multi_noalloc.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.POINTER(Parm),
                                                      ctypes.c_int, vp]
        L.ocfd_rk_stage_multi.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_advance_multi.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_advance_chem.argtypes = [vp, vp, vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(Parm), ctypes.c_int, ctypes.c_int, vp]
        L.ocfd_init_species_state.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.POINTER(Parm)]
        L.ocfd_init_species_uniform.argtypes = [vp, dp, dp, vp, vp]
        L.ocfd_level_set_inflow_species.argtypes = [vp, dp]
        L.ocfd_ctl_start_chem.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, ctypes.POINTER(Parm), vp]
        L.ocfd_step_device_chem.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Parm),
                                            ctypes.c_int, ctypes.c_int, vp]
        L.ocfd_graph_capture_steps_chem.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(Parm), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    vp, ctypes.POINTER(vp)]
        L.ocfd_tag_dengrad.argtypes = [vp, vp, ctypes.c_double, vp, vp]
        L.ocfd_level_set_march_chunk.argtypes = [vp, ctypes.c_int]
        L.ocfd_level_march_chunk.argtypes = [vp]
        L.ocfd_level_autotune.argtypes = [vp, vp, ctypes.POINTER(Parm), ctypes.c_int, ctypes.c_int, ip, dp, vp]
        L.ocfd_cluster_tags.argtypes = [ip, ip, ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(ctypes.c_ubyte),
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ip, ip]
        L.ocfd_level_set_planar.argtypes = [vp, ctypes.c_int]
        L.ocfd_tag_dengrad_nest.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, vp, vp]
        L.ocfd_cluster_box_tags.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(ctypes.c_ubyte), ctypes.c_int, ip,
                                            ip, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ip, ip]
        L.ocfd_reactor_adaptive_begin.argtypes = [ctypes.c_long, vp, ctypes.c_double, vp, vp, vp]
        L.ocfd_reactor_advance_rosenbrock.argtypes = [ctypes.c_long, vp, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp, vp, vp]
        L.ocfd_reactor_rosenbrock_begin.argtypes = [ctypes.c_long, vp, ctypes.c_double, vp, vp, vp]
        L.ocfd_reactor_rosenbrock_chunk.argtypes = [ctypes.c_long, vp, vp, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.ocfd_reactor_adaptive_chunk.argtypes = [ctypes.c_long, vp, vp, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.ocfd_project_boxes.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip, ctypes.c_int, ctypes.c_int, ip, ip]
        L.ocfd_mask_grow.argtypes = [ctypes.POINTER(ctypes.c_ubyte), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ip, ctypes.c_int]
        L.ocfd_lincomb.argtypes = [ctypes.c_long, ctypes.c_double, vp, ctypes.c_double, vp, vp, vp]
        L.ocfd_cf_fill_fine.argtypes = [vp, vp, vp, vp]
        L.ocfd_cf_interp_fine.argtypes = [vp, vp, vp, vp, vp]
        L.ocfd_copy_level.argtypes = [vp, vp, vp, vp, vp]
        L.ocfd_cf_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
        L.ocfd_cf_destroy.argtypes = [vp]
        L.ocfd_cf_nfaces.argtypes = [vp]
        L.ocfd_cf_faces.argtypes = [vp, ip]
        L.ocfd_cf_reset.argtypes = [vp, vp]
        L.ocfd_cf_add.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Parm), ctypes.c_int, vp]
        L.ocfd_cf_registers.argtypes = [vp, dp, dp]
        L.ocfd_compute_dsdt_fr.argtypes = [vp, vp, vp, ctypes.c_double, vp, vp, ctypes.POINTER(Parm), ctypes.c_int,
                                           vp]
        L.ocfd_cf_reflux.argtypes = [vp, vp, vp]
        L.ocfd_cf_avg_down.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        L.ocfd_plot_vars.argtypes = [vp, vp, vp, ctypes.POINTER(Parm), vp]
        L.ocfd_checkpoint_write_level.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, vp, ctypes.c_int, vp]
        L.ocfd_checkpoint_read_level.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, vp, ctypes.c_int, vp]
        L.ocfd_plotfile_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.c_double, ip, ctypes.POINTER(Parm), vp]
        L.ocfd_plotfile_compare.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ip, ctypes.c_char_p, ctypes.c_long]
        L.ocfd_amr_advance_level.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.POINTER(Parm), ctypes.c_int, ctypes.c_int, vp]
        L.ocfd_amr_fill_level.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, ctypes.c_double, vp]
        L.ocfd_amr_estdt.argtypes = [ctypes.c_int, vp, vp, ctypes.c_double, vp, vp, vp]
        L.ocfd_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().ocfd_last_error().decode()
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def default_parm(**kw):
    p = Parm()
    for k, v in kw.items():
        setattr(p, k, v)
    if "eos_gamma" not in kw:
        p.eos_gamma = 0.0
    check(lib().ocfd_parm_initialize(ctypes.byref(p)), "ocfd_parm_initialize")
    for k, v in kw.items():
        setattr(p, k, v)
    check(lib().ocfd_parm_initialize(ctypes.byref(p)), "ocfd_parm_initialize")
    return p


def make_geom(n_cell, prob_lo, prob_hi, periodic=(0, 0, 0), lo_bc=(2, 2, 2), hi_bc=(2, 2, 2), probtype=0,
              left_state=None, right_state=None):
    g = Geom()
    for d in range(3):
        g.dom_lo[d] = 0
        g.dom_hi[d] = n_cell[d] - 1
        g.prob_lo[d] = prob_lo[d]
        g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d]
        g.periodic[d] = periodic[d]
        g.lo_bc[d] = lo_bc[d]
        g.hi_bc[d] = hi_bc[d]
    g.probtype = probtype
    if left_state is not None:
        for n in range(5):
            g.left_state[n] = left_state[n]
            g.right_state[n] = right_state[n]
    return g


def eb_geometry(spec, geom, lo, hi, ng=NGROW):
    """Cut-cell fields of box [lo,hi] grown by ng (host computation, no GPU):
    dict of numpy arrays shaped (nz, ny, nx) (fc*: (2, nz, ny, nx), bcent: (3, ...))."""
    shape = tuple(hi[d] - lo[d] + 1 + 2 * ng for d in (2, 1, 0))
    out = {"flag": np.zeros(shape, dtype=np.uint32)}
    for k in ("vfrac", "apx", "apy", "apz"):
        out[k] = np.zeros(shape)
    for k in ("fcx", "fcy", "fcz"):
        out[k] = np.zeros((2,) + shape)
    out["bcent"] = np.zeros((3,) + shape)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    check(lib().ocfd_eb_geometry(ctypes.byref(spec), ctypes.byref(geom), (ctypes.c_int * 3)(*lo),
                                 (ctypes.c_int * 3)(*hi), ng,
                                 out["flag"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                 *(dp(out[k]) for k in ("vfrac", "apx", "apy", "apz", "fcx", "fcy", "fcz", "bcent"))),
          "ocfd_eb_geometry")
    return out


def inflow_state(p=-1.0, rho=-1.0, u=-1.0):
    st = (ctypes.c_double * 5)()
    check(lib().ocfd_inflow_state(p, rho, u, st), "ocfd_inflow_state")
    return list(st)


def dmr_states():
    left = (ctypes.c_double * 5)()
    right = (ctypes.c_double * 5)()
    check(lib().ocfd_dmr_states(left, right), "ocfd_dmr_states")
    return list(left), list(right)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


_INT_P = ctypes.POINTER(ctypes.c_int)


def box_corners(boxes):
    """(n, lo, hi): the boxes' lower and upper corners as int* arguments of the
    C ABI, 3 ints per box (each pointer keeps its array alive). One pass over
    the Python tuples, without a ctypes element per integer."""
    n = len(boxes)
    a = np.fromiter(itertools.chain.from_iterable(itertools.chain.from_iterable(boxes)), np.int32, 6 * n)
    a = a.reshape(n, 2, 3)
    lo = np.ascontiguousarray(a[:, 0]) if n else np.zeros(3, np.int32)
    hi = np.ascontiguousarray(a[:, 1]) if n else np.zeros(3, np.int32)
    return n, lo.ctypes.data_as(_INT_P), hi.ctypes.data_as(_INT_P)


def box_buffer(cap):
    """(array, int*): room for `cap` boxes written by the C ABI (6 ints each)."""
    out = np.empty(6 * max(1, cap), np.int32)
    return out, out.ctypes.data_as(_INT_P)


def boxes_from_buffer(out, nb):
    """The first `nb` boxes of a box buffer as ((lo), (hi)) tuples."""
    return [(tuple(l), tuple(h)) for l, h in out[:6 * nb].reshape(nb, 2, 3).tolist()]


class Level:
    """One AMR level's box list + geometry (drop-in for the BoxArray/Geometry
    an EBR level advances, Source/EBR.cpp:43-65)."""

    def __init__(self, boxes, geom, eb=None):
        self.boxes = [((a, b, c), (d, e, f)) for (a, b, c), (d, e, f) in boxes]
        self.geom = geom
        self.eb = eb
        n, lo, hi = box_corners(self.boxes)
        h = ctypes.c_void_p()
        check(lib().ocfd_level_create_eb(n, lo, hi, ctypes.byref(geom), ctypes.byref(eb) if eb is not None else None,
                                         ctypes.byref(h)), "ocfd_level_create_eb")
        self.handle = h

    def set_planar(self, on=True):
        """Fill only one z ghost layer (a z-uniform level stepped with PLANAR);
        stages without PLANAR are then refused."""
        check(lib().ocfd_level_set_planar(self.handle, int(bool(on))), "ocfd_level_set_planar")
        return self

    def box_types(self):
        """Per box: 0 regular, 1 cut (embedded-boundary path), 2 covered."""
        t = (ctypes.c_int * self.nbox)()
        check(lib().ocfd_level_box_types(self.handle, t), "ocfd_level_box_types")
        return list(t)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.ocfd_level_destroy(h)
            self.handle = None

    @property
    def nbox(self):
        return len(self.boxes)

    def device_bytes(self):
        """Device memory the level holds itself (not the caller's MultiFabs)."""
        return lib().ocfd_level_device_bytes(self.handle)

    def mf_size(self, ngrow=NGROW, ncomp=NCONS):
        return lib().ocfd_level_mf_size(self.handle, ngrow, ncomp)

    def fab_offset(self, b, ngrow=NGROW, ncomp=NCONS):
        return lib().ocfd_level_fab_offset(self.handle, b, ngrow, ncomp)

    def fab_shape(self, b, ngrow=NGROW, ncomp=NCONS):
        lo, hi = self.boxes[b]
        return (ncomp,) + tuple(hi[d] - lo[d] + 1 + 2 * ngrow for d in (2, 1, 0))

    def new_mf(self, ngrow=NGROW, ncomp=NCONS, device="cuda", zero=True):
        """MultiFab storage of the level; zero=False leaves it uninitialised
        (for buffers whose cells are written before they are read)."""
        import torch
        n = self.mf_size(ngrow, ncomp)
        if zero:
            return torch.zeros(n, dtype=torch.float64, device=device)
        return torch.empty(n, dtype=torch.float64, device=device)

    def fab_view(self, mf, b, ngrow=NGROW, ncomp=NCONS):
        """(ncomp, nz, ny, nx) view of box b's FAB inside MultiFab storage mf."""
        off = self.fab_offset(b, ngrow, ncomp)
        shape = self.fab_shape(b, ngrow, ncomp)
        return mf[off:off + int(np.prod(shape))].view(shape)

    # ---- entry points -------------------------------------------------
    def compute_dsdt(self, S, dSdt, parm, flags=0, dt=0.0, stream=None):
        check(lib().ocfd_compute_dsdt(self.handle, _ptr(S), _ptr(dSdt), dt, ctypes.byref(parm), flags,
                                      _stream(stream)), "ocfd_compute_dsdt")

    def compute_dsdt_fr(self, S, dSdt, parm, flags=0, dt=0.0, fine=None, current=None, stream=None):
        """compute_dsdt, then dt x the interface fluxes added to the register
        `fine` (this level its coarse side) and `current` (its fine side);
        registers are amr.FluxRegister objects or None."""
        h = lambda r: None if r is None else r.handle
        check(lib().ocfd_compute_dsdt_fr(self.handle, _ptr(S), _ptr(dSdt), dt, h(fine), h(current),
                                         ctypes.byref(parm), flags, _stream(stream)), "ocfd_compute_dsdt_fr")

    def rk_stage(self, S_cur, S_old, S_new, c_cur, c_old, c_dt, parm, flags=0, stream=None):
        check(lib().ocfd_rk_stage(self.handle, _ptr(S_cur), _ptr(S_old), _ptr(S_new), c_cur, c_old, c_dt,
                                  ctypes.byref(parm), flags, _stream(stream)), "ocfd_rk_stage")

    def march_chunk(self, ch=None):
        """Cells per march chunk of the y/z sweeps (set it when ch is given)."""
        if ch is not None:
            check(lib().ocfd_level_set_march_chunk(self.handle, ch), "ocfd_level_set_march_chunk")
        return lib().ocfd_level_march_chunk(self.handle)

    def autotune(self, S, parm, flags=0, reps=5, stream=None):
        """Time the march tile shapes on state S (ghosts filled) and keep the
        fastest; returns (chunk, {chunk: ms})."""
        ch = ctypes.c_int()
        ms = (ctypes.c_double * 6)()
        check(lib().ocfd_level_autotune(self.handle, _ptr(S), ctypes.byref(parm), flags, reps, ctypes.byref(ch), ms,
                                        _stream(stream)), "ocfd_level_autotune")
        return ch.value, dict(zip((2, 4, 8, 16, 32, 64), list(ms)))

    def fill_boundary(self, S, time, stream=None):
        check(lib().ocfd_fill_boundary(self.handle, _ptr(S), time, _stream(stream)), "ocfd_fill_boundary")

    def estdt(self, S, cfl, parm, stream=None):
        out = ctypes.c_double()
        check(lib().ocfd_estdt(self.handle, _ptr(S), cfl, ctypes.byref(parm), ctypes.byref(out), _stream(stream)),
              "ocfd_estdt")
        return out.value

    def advance(self, A, B, C, time, dt, scheme, parm, flags=0, stream=None):
        check(lib().ocfd_advance(self.handle, _ptr(A), _ptr(B), _ptr(C), time, dt, scheme, ctypes.byref(parm),
                                 flags, _stream(stream)), "ocfd_advance")

    def profile_stage(self, S_cur, S_old, S_new, c_cur, c_old, c_dt, parm, flags=0, reps=5, stream=None):
        """Average ms of (x sweep, y sweep, z sweep + update, ghost fill, primitives)."""
        out = (ctypes.c_double * 5)()
        check(lib().ocfd_profile_stage(self.handle, _ptr(S_cur), _ptr(S_old), _ptr(S_new), c_cur, c_old, c_dt,
                                       ctypes.byref(parm), flags, reps, out, _stream(stream)), "ocfd_profile_stage")
        return list(out)

    def chem_advance(self, S, Spec, dt, parm, stream=None):
        check(lib().ocfd_chem_advance(self.handle, _ptr(S), _ptr(Spec), dt, ctypes.byref(parm), _stream(stream)),
              "ocfd_chem_advance")

    def species_renorm(self, S, Spec, stream=None):
        check(lib().ocfd_species_renorm(self.handle, _ptr(S), _ptr(Spec), _stream(stream)), "ocfd_species_renorm")

    def init_state(self, kind, S, parm, seed=0):
        check(lib().ocfd_init_state(self.handle, kind, _ptr(S), ctypes.byref(parm), seed), "ocfd_init_state")

    # ---- multispecies flow ----------------------------------------------
    def new_spec_mf(self, ngrow=NGROW, device="cuda"):
        return self.new_mf(ngrow, NSPEC, device)

    def fill_boundary_spec(self, Spec, time, stream=None):
        check(lib().ocfd_fill_boundary_spec(self.handle, _ptr(Spec), time, _stream(stream)), "ocfd_fill_boundary_spec")

    def compute_dsdt_multi(self, S, Spec, dSdt, dSdt_spec, parm, flags=0, dt=0.0, stream=None):
        check(lib().ocfd_compute_dsdt_multi(self.handle, _ptr(S), _ptr(Spec), _ptr(dSdt), _ptr(dSdt_spec), dt,
                                            ctypes.byref(parm), flags, _stream(stream)), "ocfd_compute_dsdt_multi")

    def advance_multi(self, A, As, B, Bs, time, dt, scheme, parm, flags=0, stream=None):
        check(lib().ocfd_advance_multi(self.handle, _ptr(A), _ptr(As), _ptr(B), _ptr(Bs), time, dt, scheme,
                                       ctypes.byref(parm), flags, _stream(stream)), "ocfd_advance_multi")

    def advance_chem(self, A, As, B, Bs, time, dt, parm, level=0, scheme=1, flags=0, renorm=True, stream=None):
        check(lib().ocfd_advance_chem(self.handle, _ptr(A), _ptr(As), _ptr(B), _ptr(Bs), time, dt, level, scheme,
                                      ctypes.byref(parm), flags, int(renorm), _stream(stream)), "ocfd_advance_chem")

    def set_inflow_species(self, spec):
        """Partial densities the species ghost fill writes on ext_dir sides (None clears)."""
        dp = ctypes.POINTER(ctypes.c_double)
        self._spec_in = None if spec is None else np.ascontiguousarray(spec, dtype=np.float64)
        check(lib().ocfd_level_set_inflow_species(self.handle, None if spec is None else
                                                  self._spec_in.ctypes.data_as(dp)), "ocfd_level_set_inflow_species")

    def init_species_uniform(self, state, spec, S, Spec):
        """Every cell (ghosts included) of S and Spec set to `state` (5) and `spec` (9)."""
        st = np.ascontiguousarray(state, dtype=np.float64)
        sp = np.ascontiguousarray(spec, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().ocfd_init_species_uniform(self.handle, st.ctypes.data_as(dp), sp.ctypes.data_as(dp), _ptr(S),
                                              _ptr(Spec)), "ocfd_init_species_uniform")

    def ctl_start_chem(self, A, ctl, parm, cfl, time=0.0, stop_time=-1.0, max_steps=-1, stream=None):
        check(lib().ocfd_ctl_start_chem(self.handle, _ptr(A), _ptr(ctl), time, cfl, stop_time, max_steps,
                                        ctypes.byref(parm), _stream(stream)), "ocfd_ctl_start_chem")

    def step_device_chem(self, A, As, B, Bs, ctl, parm, scheme=1, level=0, flags=0, renorm=True, stream=None):
        check(lib().ocfd_step_device_chem(self.handle, _ptr(A), _ptr(As), _ptr(B), _ptr(Bs), _ptr(ctl), scheme,
                                          level, ctypes.byref(parm), flags, int(renorm), _stream(stream)),
              "ocfd_step_device_chem")

    def capture_steps_chem(self, A, As, B, Bs, ctl, parm, nsteps, scheme=1, level=0, flags=0, renorm=True,
                           stream=None):
        """hipGraph of nsteps device-driven multispecies steps (stream must not be None)."""
        return StepGraph(self, A, B, None, ctl, parm, nsteps, scheme, flags, stream,
                         chem=(As, Bs, level, int(renorm)))

    def init_species_state(self, kind, S, Spec, parm):
        check(lib().ocfd_init_species_state(self.handle, kind, _ptr(S), _ptr(Spec), ctypes.byref(parm)),
              "ocfd_init_species_state")

    # ---- device-driven time steps ---------------------------------------
    def ctl_start(self, A, ctl, parm, cfl, time=0.0, stop_time=-1.0, max_steps=-1, stream=None):
        """Initial dt on the device into ctl (a CTL_BYTES-byte device buffer)."""
        check(lib().ocfd_ctl_start(self.handle, _ptr(A), _ptr(ctl), time, cfl, stop_time, max_steps,
                                   ctypes.byref(parm), _stream(stream)), "ocfd_ctl_start")

    def step_device(self, A, B, C, ctl, parm, scheme=2, flags=0, stream=None):
        check(lib().ocfd_step_device(self.handle, _ptr(A), _ptr(B), _ptr(C), _ptr(ctl), scheme, ctypes.byref(parm),
                                     flags, _stream(stream)), "ocfd_step_device")

    def capture_steps(self, A, B, C, ctl, parm, nsteps, scheme=2, flags=0, stream=None):
        """A StepGraph of nsteps device-driven steps, captured on `stream`
        (a non-default torch stream)."""
        return StepGraph(self, A, B, C, ctl, parm, nsteps, scheme, flags, stream)


CHEM_ESTDT = 1e-8   # the CHEM build's ebr_estdt (Source/Advance.cpp:54-57)


def run_chem(level, A, As, B, Bs, parm, cfl, nsteps=-1, stop_time=-1.0, scheme=1, flags=0, lev=0, time=0.0,
             stream=None):
    """Time loop of a multispecies level: dt = cfl * 1e-8 (the CHEM build's
    estimate), computeNewDt's change_max and stop_time clip
    (Source/EBR.cpp:231-306), one Strang-split EBR::advance plus the species
    renormalisation per step."""
    dt = cfl * CHEM_ESTDT
    if stop_time >= 0.0 and time + dt > stop_time - 0.001 * dt:
        dt = stop_time - time
    steps = 0
    while (nsteps < 0 or steps < nsteps) and (stop_time < 0.0 or time < stop_time):
        level.advance_chem(A, As, B, Bs, time, dt, parm, lev, scheme, flags, True, stream)
        time += dt
        steps += 1
        dt_new = min(cfl * CHEM_ESTDT, 1.1 * dt)
        if stop_time >= 0.0 and time + dt_new > stop_time - 0.001 * dt_new:
            dt_new = stop_time - time
        dt = dt_new
    return steps, time, dt


def level_error(level, got, ref, ncomp=NCONS):
    """Whole-level relative (L1, Linf) error of state `got` against `ref`
    (two MultiFabs of `level`, 4 ghosts) per conserved variable, over valid
    cells, computed on the device: sum/max of |got - ref| over sum/max of
    |ref|.  The momentum components share the momentum vector's norm."""
    import torch
    num1 = torch.zeros(ncomp, dtype=torch.float64, device=got.device)
    numi = torch.zeros_like(num1); den1 = torch.zeros_like(num1); deni = torch.zeros_like(num1)
    for b in range(level.nbox):
        a = level.fab_view(got, b, NGROW, ncomp)[:, 4:-4, 4:-4, 4:-4].reshape(ncomp, -1)
        r = level.fab_view(ref, b, NGROW, ncomp)[:, 4:-4, 4:-4, 4:-4].reshape(ncomp, -1)
        d = (a - r).abs()
        num1 += d.sum(1); numi = torch.maximum(numi, d.amax(1))
        den1 += r.abs().sum(1); deni = torch.maximum(deni, r.abs().amax(1))
    num1, numi, den1, deni = (x.cpu().numpy() for x in (num1, numi, den1, deni))
    if ncomp >= 4:
        den1[1:4] = den1[1:4].sum()
        deni[1:4] = deni[1:4].max()
    den1 = np.maximum(den1, 1e-300); deni = np.maximum(deni, 1e-300)
    return [(float(x), float(y)) for x, y in zip(num1 / den1, numi / deni)]


def cached_device_bytes():
    """Device bytes the library's pool holds free for reuse."""
    return lib().ocfd_cached_device_bytes()


def release_cached_memory():
    """Return the pool's free device blocks to HIP."""
    check(lib().ocfd_release_cached_memory(), "ocfd_release_cached_memory")


def new_ctl(device="cuda"):
    import torch
    return torch.zeros(CTL_BYTES, dtype=torch.uint8, device=device)


def read_ctl(ctl):
    """Host copy of a device StepCtl (synchronises with its stream)."""
    return StepCtl.from_buffer_copy(bytes(ctl.cpu().numpy().tobytes()))


class StepGraph:
    """hipGraph of device-driven time steps; replayed with launch()."""

    def __init__(self, level, A, B, C, ctl, parm, nsteps, scheme=2, flags=0, stream=None, chem=None):
        if stream is None:
            raise ValueError("graph capture needs a non-default stream")
        self.level = level                     # keeps the level's work-lists alive
        self.keep = (A, B, C, ctl, chem)      # and the buffers the graph writes
        self.parm = parm
        h = ctypes.c_void_p()
        if chem is None:
            check(lib().ocfd_graph_capture_steps(level.handle, _ptr(A), _ptr(B), _ptr(C), _ptr(ctl), scheme,
                                                 ctypes.byref(parm), flags, nsteps, _stream(stream),
                                                 ctypes.byref(h)), "ocfd_graph_capture_steps")
        else:
            As, Bs, lev, renorm = chem
            check(lib().ocfd_graph_capture_steps_chem(level.handle, _ptr(A), _ptr(As), _ptr(B), _ptr(Bs), _ptr(ctl),
                                                      scheme, lev, ctypes.byref(parm), flags, renorm, nsteps,
                                                      _stream(stream), ctypes.byref(h)),
                  "ocfd_graph_capture_steps_chem")
        self.handle = h
        self.nsteps = nsteps

    def launch(self, stream=None):
        check(lib().ocfd_graph_launch(self.handle, _stream(stream)), "ocfd_graph_launch")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.ocfd_graph_destroy(h)
            self.handle = None


def run(level, A, B, C, parm, cfl, scheme=2, flags=0, stop_time=-1.0, nsteps=-1, time=0.0, dt=0.0, stream=None):
    """Single-level time loop of the reference driver: initial dt =
    estTimeStep (Source/EBR.cpp:189-229,561-563), then per step computeNewDt
    with change_max 1.1 and the stop_time clip (Source/EBR.cpp:231-306)."""
    if dt <= 0.0:
        dt = level.estdt(A, cfl, parm, stream)
        eps = 0.001 * dt
        if stop_time >= 0.0 and time + dt > stop_time - eps:
            dt = stop_time - time
    steps = 0
    while (nsteps < 0 or steps < nsteps) and (stop_time < 0.0 or time < stop_time):
        level.advance(A, B, C, time, dt, scheme, parm, flags, stream)
        time += dt
        steps += 1
        est = level.estdt(A, cfl, parm, stream)
        dt_new = min(est, 1.1 * dt)
        eps = 0.001 * dt_new
        if stop_time >= 0.0 and time + dt_new > stop_time - eps:
            dt_new = stop_time - time
        dt = dt_new
    return steps, time, dt


# ---------------------------------------------------------------------------
# batched constant-volume reactors (0D ignition)
# ---------------------------------------------------------------------------
REACTOR_NCOMP = 11   # rows: rho, E, 9 partial densities


class Reactors:
    """A device-resident batch of n constant-volume H2/air reactors."""

    def __init__(self, T0, p0, phi, stream=None):
        import torch
        self.T0 = torch.as_tensor(T0, dtype=torch.float64, device="cuda").contiguous()
        p0 = torch.as_tensor(p0, dtype=torch.float64, device="cuda").contiguous()
        phi = torch.as_tensor(phi, dtype=torch.float64, device="cuda").contiguous()
        self.n = int(self.T0.numel())
        self.state = torch.zeros(REACTOR_NCOMP * self.n, dtype=torch.float64, device="cuda")
        self.steps = torch.zeros(self.n, dtype=torch.int32, device="cuda")
        self.rejects = torch.zeros(self.n, dtype=torch.int32, device="cuda")
        self.work = torch.zeros(1, dtype=torch.int64, device="cuda")
        check(lib().ocfd_reactor_init(self.n, _ptr(self.T0), _ptr(p0), _ptr(phi), _ptr(self.state),
                                      _stream(stream)), "ocfd_reactor_init")

    def rows(self):
        return self.state.view(REACTOR_NCOMP, self.n)

    def advance_fixed(self, dt, nsteps, stream=None):
        check(lib().ocfd_reactor_advance_fixed(self.n, _ptr(self.state), dt, nsteps, _stream(stream)),
              "ocfd_reactor_advance_fixed")

    def advance_adaptive(self, t_end, h0=1e-8, rtol=1e-4, atol=1e-9, max_steps=200000, stream=None):
        check(lib().ocfd_reactor_advance_adaptive(self.n, _ptr(self.state), t_end, h0, rtol, atol, max_steps,
                                                  _ptr(self.steps), _ptr(self.rejects), _ptr(self.work),
                                                  _stream(stream)), "ocfd_reactor_advance_adaptive")

    def advance_rosenbrock(self, t_end, h0=1e-8, rtol=1e-4, atol=1e-9, max_steps=200000, stream=None):
        """RODAS3 integration of every reactor to t_end."""
        check(lib().ocfd_reactor_advance_rosenbrock(self.n, _ptr(self.state), t_end, h0, rtol, atol, max_steps,
                                                    _ptr(self.steps), _ptr(self.rejects), _ptr(self.work),
                                                    _stream(stream)), "ocfd_reactor_advance_rosenbrock")

    def begin_chunks(self, h0=1e-8, stream=None, method="adaptive"):
        """Resumable integration: per-reactor time 0 and step h0; method
        "adaptive" (step doubling) or "rosenbrock" (RODAS3)."""
        import torch
        if method not in ("adaptive", "rosenbrock"):
            raise ValueError("method must be 'adaptive' or 'rosenbrock'")
        self.method = method
        self.th = torch.empty((3 if method == "rosenbrock" else 2) * self.n, dtype=torch.float64, device="cuda")
        self.remaining = torch.zeros(1, dtype=torch.int32, device="cuda")
        begin = lib().ocfd_reactor_rosenbrock_begin if method == "rosenbrock" else lib().ocfd_reactor_adaptive_begin
        check(begin(self.n, _ptr(self.th), h0, _ptr(self.steps), _ptr(self.rejects), _stream(stream)),
              "ocfd_reactor_" + method + "_begin")

    def chunk(self, t_end, budget=16, rtol=1e-4, atol=1e-9, max_steps=200000, stream=None):
        """At most `budget` substep attempts per reactor; self.remaining (device)
        then counts the reactors short of t_end."""
        fn = lib().ocfd_reactor_rosenbrock_chunk if self.method == "rosenbrock" else lib().ocfd_reactor_adaptive_chunk
        check(fn(self.n, _ptr(self.state), _ptr(self.th), t_end, rtol, atol, max_steps, budget, _ptr(self.steps),
                 _ptr(self.rejects), _ptr(self.remaining), _stream(stream)), "ocfd_reactor_" + self.method + "_chunk")

    def temperature(self, stream=None):
        import torch
        T = torch.empty(self.n, dtype=torch.float64, device="cuda")
        check(lib().ocfd_reactor_temperature(self.n, _ptr(self.state), _ptr(T), _stream(stream)),
              "ocfd_reactor_temperature")
        return T


def premixed_state(T0, p0, phi, u0=0.0):
    """Conserved state (rho, rho u, rho v, rho w, E) and the 9 partial
    densities of premixed H2/air (X(H2):X(O2):X(N2) = 2phi:1:3.76) at T0 [K],
    p0 [Pa], moving at u0 [m/s] along x; E is the sensible energy density plus
    the kinetic energy (the library's reactor initialisation, CKXTY/CKRHOY/
    CKUBMS of Source/CHEM/LiDryer.H)."""
    R = Reactors([T0], [p0], [phi])
    rows = R.rows()[:, 0].cpu().numpy()
    rho, E, rhoi = rows[0], rows[1], rows[2:].copy()
    state = np.array([rho, rho * u0, 0.0, 0.0, E + 0.5 * rho * u0 * u0])
    return state, rhoi


def ignition_batch(n, seed=7):
    """Seeded 0D ignition inputs: T0 ~ U[900,1500] K, p0 ~ U[1,10] atm, phi ~ U[0.5,2]."""
    rng = np.random.default_rng(seed)
    T0 = rng.uniform(900.0, 1500.0, n)
    p0 = rng.uniform(1.0, 10.0, n) * 101325.0
    phi = rng.uniform(0.5, 2.0, n)
    return T0, p0, phi
