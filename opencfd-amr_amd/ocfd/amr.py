"""Two-level AMR time stepping on the device (refinement ratio 2, subcycled).

The driver follows the AmrLevel/Amr semantics the reference runs under
(Source/EBR.cpp, Source/Advance.cpp):

* initial grids: the coarse level is initialised, its cells are tagged with
  tag_dengrad (Source/Tagging.H:8-36, threshold ebr.refine_dengrad), the tags
  are clustered into fine boxes, the fine level is initialised from the case
  and averaged down (EBR::post_init, Source/EBR.cpp:437-448);
* every coarse step: dt from EBR::computeInitialDt (first step) or
  EBR::computeNewDt (change_max 1.1, subcycling factor 2, stop-time clip;
  Source/EBR.cpp:189-306); a regrid every amr.regrid_int coarse steps; the
  coarse advance with its flux-register adds; two fine steps of dt/2 whose
  stage ghost cells come from the fine level and from the coarse level
  interpolated to the stage time; then reflux and average-down
  (EBR::post_timestep / EBR::avgDown, Source/EBR.cpp:313-330,535-559).

Stage times and flux-register weights per scheme (Source/Advance.cpp:117-148):
Euler (t; dt), RK2 (t, t+dt; dt/2, dt/2), SSP-RK3 (t, t+dt, t+dt/2; dt/6,
dt/6, 2dt/3).  All device work is issued on one stream; the host only reads
the dt reductions and the tags of a regrid.
"""
import ctypes
import importlib

import numpy as np

from . import PLANAR, Geom, Level, NCONS, _ptr, _stream, check, lib

CHANGE_MAX = 1.1
RATIO = 2
_pkg = importlib.import_module(__package__)


def refine_geom(g):
    """Geometry of the level refined by 2 (same physical box and BCs)."""
    f = Geom()
    ctypes.memmove(ctypes.byref(f), ctypes.byref(g), ctypes.sizeof(Geom))
    for d in range(3):
        f.dom_lo[d] = RATIO * g.dom_lo[d]
        f.dom_hi[d] = RATIO * g.dom_hi[d] + RATIO - 1
        f.dx[d] = g.dx[d] / RATIO
    return f


def refine_boxes(boxes):
    return [(tuple(RATIO * v for v in lo), tuple(RATIO * v + RATIO - 1 for v in hi)) for lo, hi in boxes]


def lincomb(a, x, b, y, out, stream=None):
    check(lib().ocfd_lincomb(out.numel(), a, _ptr(x), b, _ptr(y), _ptr(out), _stream(stream)), "ocfd_lincomb")


def tag_dengrad(level, S, threshold, stream=None):
    """One byte per valid cell (the level's 0-ghost layout), 1 = tagged."""
    import torch
    tags = torch.zeros(level.mf_size(0, 1), dtype=torch.uint8, device=S.device)
    check(lib().ocfd_tag_dengrad(level.handle, _ptr(S), threshold, _ptr(tags), _stream(stream)), "ocfd_tag_dengrad")
    return tags


def dense_tags(level, tags):
    """Tags of a level (host array in the 0-ghost layout) as a dense (nz, ny, nx)
    array over the level's domain."""
    g = level.geom
    n = [g.dom_hi[d] - g.dom_lo[d] + 1 for d in range(3)]
    out = np.zeros((n[2], n[1], n[0]), dtype=np.uint8)
    for b, (lo, hi) in enumerate(level.boxes):
        off = level.fab_offset(b, 0, 1)
        shp = tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
        cnt = int(np.prod(shp))
        out[lo[2] - g.dom_lo[2]:hi[2] - g.dom_lo[2] + 1, lo[1] - g.dom_lo[1]:hi[1] - g.dom_lo[1] + 1,
            lo[0] - g.dom_lo[0]:hi[0] - g.dom_lo[0] + 1] = tags[off:off + cnt].reshape(shp)
    return out


def cluster_tags(dense, lo, block=1, grid_eff=0.7, max_size=32):
    """Boxes (in the tags' index space, domain lower corner `lo`) covering the
    tagged cells of a dense (nz, ny, nx) array."""
    dense = np.ascontiguousarray(dense, dtype=np.uint8)
    nz, ny, nx = dense.shape
    lo_c = (ctypes.c_int * 3)(*lo)
    hi_c = (ctypes.c_int * 3)(lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
    cap = max(16, dense.size)
    out = (ctypes.c_int * (6 * cap))()
    nb = ctypes.c_int()
    check(lib().ocfd_cluster_tags(lo_c, hi_c, dense.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), block, grid_eff,
                                  max_size, cap, out, ctypes.byref(nb)), "ocfd_cluster_tags")
    return [(tuple(out[6 * b:6 * b + 3]), tuple(out[6 * b + 3:6 * b + 6])) for b in range(nb.value)]


def fill_from_coarse(fine, S_fine, coarse, S_coarse, stream=None):
    check(lib().ocfd_fill_from_coarse(fine.handle, _ptr(S_fine), coarse.handle, _ptr(S_coarse), _stream(stream)),
          "ocfd_fill_from_coarse")


def interp_from_coarse(fine, S_fine, coarse, S_coarse, old_fine=None, stream=None):
    check(lib().ocfd_interp_from_coarse(fine.handle, _ptr(S_fine), coarse.handle, _ptr(S_coarse),
                                        old_fine.handle if old_fine is not None else None, _stream(stream)),
          "ocfd_interp_from_coarse")


def copy_level(dst, S_dst, src, S_src, stream=None):
    check(lib().ocfd_copy_level(dst.handle, _ptr(S_dst), src.handle, _ptr(S_src), _stream(stream)), "ocfd_copy_level")


class AmrBufs(ctypes.Structure):
    """struct ocfd_amr_bufs: the state buffers of one two-level advance."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("c_old", "c_new", "c_s1", "c_s2", "c_t", "f_state", "f_s1", "f_s2")]


class FluxRegister:
    """The coarse/fine interface of one fine level (EBR's flux_reg)."""

    def __init__(self, coarse, fine):
        self.coarse, self.fine = coarse, fine
        h = ctypes.c_void_p()
        check(lib().ocfd_cf_create(coarse.handle, fine.handle, ctypes.byref(h)), "ocfd_cf_create")
        self.handle = h
        self.nfaces = lib().ocfd_cf_nfaces(h)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _pkg is not None and _pkg._lib is not None:
            _pkg._lib.ocfd_cf_destroy(h)
            self.handle = None

    def faces(self):
        """(nface, 6) array: coarse box, i, j, k, dir, high-face flag."""
        out = (ctypes.c_int * max(1, 6 * self.nfaces))()
        check(lib().ocfd_cf_faces(self.handle, out), "ocfd_cf_faces")
        return np.frombuffer(out, dtype=np.int32)[:6 * self.nfaces].reshape(self.nfaces, 6).copy()

    def reset(self, stream=None):
        check(lib().ocfd_cf_reset(self.handle, _stream(stream)), "ocfd_cf_reset")

    def add(self, fine_side, w_dt, parm, flags=0, stream=None):
        check(lib().ocfd_cf_add(self.handle, int(fine_side), w_dt, ctypes.byref(parm), flags, _stream(stream)),
              "ocfd_cf_add")

    def registers(self):
        rc = np.zeros(max(1, 5 * self.nfaces))
        rf = np.zeros(max(1, 20 * self.nfaces))
        check(lib().ocfd_cf_registers(self.handle, rc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      rf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "ocfd_cf_registers")
        return rc[:5 * self.nfaces].reshape(-1, 5), rf[:20 * self.nfaces].reshape(-1, 4, 5)

    def reflux(self, S_coarse, stream=None):
        check(lib().ocfd_cf_reflux(self.handle, _ptr(S_coarse), _stream(stream)), "ocfd_cf_reflux")

    def avg_down(self, S_fine, S_coarse, ncomp=NCONS, stream=None):
        check(lib().ocfd_cf_avg_down(self.handle, _ptr(S_fine), _ptr(S_coarse), ncomp, _stream(stream)),
              "ocfd_cf_avg_down")


class TwoLevel:
    """A coarse level covering the domain plus one fine level of ratio 2.

    coarse_boxes, geom: the base level.  kind: case initial data of
    ocfd_init_state (None: set the states yourself and call start()).
    fine_boxes: fixed fine grids (coarse-index boxes are refined), else the
    fine grids come from tagging.  regrid_int 0 keeps the grids fixed.
    blocking_factor and max_grid_size are in fine cells (amr.* inputs).
    """

    def __init__(self, coarse_boxes, geom, parm, cfl, kind=None, scheme=2, flags=0, refine_dengrad=0.05,
                 regrid_int=2, blocking_factor=2, grid_eff=0.7, max_grid_size=64, fine_boxes=None, seed=0,
                 stop_time=-1.0, do_reflux=True, stream=None):
        if scheme not in (0, 1, 2):
            raise ValueError("scheme must be 0 (Euler), 1 (RK2) or 2 (SSP-RK3)")
        if blocking_factor % RATIO or max_grid_size % blocking_factor:
            raise ValueError("blocking_factor must be a multiple of 2 and divide max_grid_size")
        self.parm, self.cfl, self.scheme, self.flags = parm, cfl, scheme, flags
        self.thr, self.regrid_int = refine_dengrad, regrid_int
        self.block, self.grid_eff, self.max_c = blocking_factor // RATIO, grid_eff, max_grid_size // RATIO
        if stream is None:
            import torch
            stream = torch.cuda.current_stream()
        self.stop_time, self.do_reflux, self.stream = stop_time, do_reflux, stream
        self.kind, self.seed = kind, seed
        self.coarse = Level(coarse_boxes, geom)
        self.gf = refine_geom(geom)
        self.S = [self.coarse.new_mf(), None]          # new state per level
        self.Oc = self.coarse.new_mf()                 # coarse old state
        self.scratch_c = [self.coarse.new_mf(), self.coarse.new_mf(), self.coarse.new_mf()]
        self.fine = None
        self.cf = None
        self.scratch_f = None
        self.time = 0.0
        self.dt_level = [0.0, 0.0]
        self.steps = 0
        self.count = 0
        self.fixed_fine = fine_boxes
        if kind is not None:
            self.coarse.init_state(kind, self.S[0], parm, seed)
            self.start()

    # -- grids ------------------------------------------------------------
    def _fine_boxes_from_tags(self):
        c = self.coarse
        c.fill_boundary(self.S[0], self.time, self.stream)
        tags = tag_dengrad(c, self.S[0], self.thr, self.stream).cpu().numpy()
        dense = dense_tags(c, tags)
        g = c.geom
        boxes = cluster_tags(dense, (g.dom_lo[0], g.dom_lo[1], g.dom_lo[2]), self.block, self.grid_eff, self.max_c)
        return refine_boxes(boxes)

    def _set_fine(self, level, S):
        self.fine = level
        self.S[1] = S
        if level is None:
            self.cf, self.scratch_f = None, None
            return
        self.scratch_f = [level.new_mf(), level.new_mf()]
        self.cf = FluxRegister(self.coarse, level)
        if (self.flags & PLANAR) and self.cf.nfaces and (self.cf.faces()[:, 4] == 2).any():
            raise ValueError("planar stepping needs fine grids that span the domain in z")

    def start(self):
        """Initial grids (bldFineLevels + post_init): fine grids from the
        coarse tags or the fixed list, the case on the fine level, average-down."""
        boxes = refine_boxes(self.fixed_fine) if self.fixed_fine is not None else self._fine_boxes_from_tags()
        if not boxes:
            self._set_fine(None, None)
            return
        f = Level(boxes, self.gf)
        Sf = f.new_mf()
        if self.kind is not None:
            f.init_state(self.kind, Sf, self.parm, self.seed)
        else:
            self.coarse.fill_boundary(self.S[0], self.time, self.stream)
            interp_from_coarse(f, Sf, self.coarse, self.S[0], None, self.stream)
        self._set_fine(f, Sf)
        self.cf.avg_down(Sf, self.S[0], NCONS, self.stream)

    def regrid(self):
        """New fine grids from the current coarse state; the new fine level takes
        the old fine level's cells where it had them and the interpolated
        coarse state elsewhere (AmrLevel::init(old))."""
        boxes = self._fine_boxes_from_tags()      # fills the coarse ghosts at self.time
        if not boxes:
            self._set_fine(None, None)
            return
        f = Level(boxes, self.gf)
        Sf = f.new_mf()
        if self.fine is not None:
            copy_level(f, Sf, self.fine, self.S[1], self.stream)
        interp_from_coarse(f, Sf, self.coarse, self.S[0], self.fine, self.stream)
        self._set_fine(f, Sf)

    # -- time step ----------------------------------------------------------
    def _estdt(self):
        est = [self.coarse.estdt(self.S[0], self.cfl, self.parm, self.stream)]
        if self.fine is not None:
            est.append(self.fine.estdt(self.S[1], self.cfl, self.parm, self.stream))
        return est

    def _compute_dt(self):
        est = self._estdt()
        if self.steps > 0:
            prev = [self.dt_level[0], self.dt_level[1] if self.dt_level[1] > 0 else self.dt_level[0] / RATIO]
            est = [min(e, CHANGE_MAX * prev[i]) for i, e in enumerate(est)]
        dt0 = est[0]
        if len(est) > 1:
            dt0 = min(dt0, RATIO * est[1])
        eps = 0.001 * dt0
        if self.stop_time >= 0.0 and self.time + dt0 > self.stop_time - eps:
            dt0 = self.stop_time - self.time
        self.dt_level = [dt0, dt0 / RATIO]

    def _bufs(self):
        b = AmrBufs()
        b.c_old, b.c_new = self.Oc.data_ptr(), self.S[0].data_ptr()
        b.c_s1, b.c_s2, b.c_t = (x.data_ptr() for x in self.scratch_c)
        if self.fine is not None:
            b.f_state = self.S[1].data_ptr()
            b.f_s1, b.f_s2 = (x.data_ptr() for x in self.scratch_f)
        return b

    def step(self):
        """One coarse time step (Amr::coarseTimeStep with two subcycled fine steps)."""
        self._compute_dt()
        if self.fixed_fine is None and self.regrid_int > 0 and self.count >= self.regrid_int:
            self.regrid()
            self.count = 0
        t, dt0 = self.time, self.dt_level[0]
        self.Oc, self.S[0] = self.S[0], self.Oc          # old <- current; the new state is written
        b = self._bufs()
        check(lib().ocfd_amr_advance(self.coarse.handle, self.cf.handle if self.cf is not None else None,
                                     ctypes.byref(b), t, dt0, self.scheme, ctypes.byref(self.parm), self.flags,
                                     int(self.do_reflux), _stream(self.stream)), "ocfd_amr_advance")
        self.time = t + dt0
        self.steps += 1
        self.count += 1
        return dt0

    def run(self, nsteps=-1):
        n = 0
        while (nsteps < 0 or n < nsteps) and (self.stop_time < 0.0 or self.time < self.stop_time):
            self.step()
            n += 1
        return n

    def fill(self):
        """Fill the ghost cells of both levels at the current time."""
        self.coarse.fill_boundary(self.S[0], self.time, self.stream)
        if self.fine is not None:
            self.fine.fill_boundary(self.S[1], self.time, self.stream)
            fill_from_coarse(self.fine, self.S[1], self.coarse, self.S[0], self.stream)
