"""AMR hierarchies on the device (any depth, refinement ratio 2, subcycled).

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
import os

import numpy as np

from . import PLANAR, Geom, Level, NCONS, _ptr, _stream, box_buffer, box_corners, boxes_from_buffer, check, lib

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
    r, e = RATIO, RATIO - 1
    return [((r * a, r * b, r * c), (r * d + e, r * f + e, r * g + e)) for (a, b, c), (d, f, g) in boxes]


def lincomb(a, x, b, y, out, stream=None):
    check(lib().ocfd_lincomb(out.numel(), a, _ptr(x), b, _ptr(y), _ptr(out), _stream(stream)), "ocfd_lincomb")


def tag_dengrad(level, S, threshold, stream=None):
    """One byte per valid cell (the level's 0-ghost layout), 1 = tagged."""
    import torch
    tags = torch.zeros(level.mf_size(0, 1), dtype=torch.uint8, device=S.device)
    check(lib().ocfd_tag_dengrad(level.handle, _ptr(S), threshold, _ptr(tags), _stream(stream)), "ocfd_tag_dengrad")
    return tags


def tag_dengrad_nest(level, S, threshold, n_proper, stream=None):
    """Device flags, one byte per valid cell (0-ghost layout): bit 0 the
    density-gradient tag, bit 1 inside the level's proper nesting region."""
    import torch
    flags = torch.empty(level.mf_size(0, 1), dtype=torch.uint8, device=S.device)
    check(lib().ocfd_tag_dengrad_nest(level.handle, _ptr(S), threshold, int(n_proper), _ptr(flags), _stream(stream)),
          "ocfd_tag_dengrad_nest")
    return flags


def cluster_box_tags(boxes, flags, extra=(), use_mask=True, block=1, grid_eff=0.7, max_size=32):
    """cluster_tags from flags in the level layout of `boxes` (host uint8,
    bit 0 tag, bit 1 allowed region) plus the cells of the `extra` boxes
    as tags; the clusters are cut back to the allowed region when use_mask."""
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    n, lo, hi = box_corners(boxes)
    ne, elo, ehi = box_corners(list(extra))
    nb = ctypes.c_int()
    cap = 4096
    while True:
        out, outp = box_buffer(cap)
        rc = lib().ocfd_cluster_box_tags(n, lo, hi, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), ne, elo, ehi,
                                         int(bool(use_mask)), block, grid_eff, max_size, cap, outp, ctypes.byref(nb))
        if rc != 0 and nb.value > cap:
            cap = nb.value
            continue
        check(rc, "ocfd_cluster_box_tags")
        break
    return boxes_from_buffer(out, nb.value)


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


def cluster_tags(dense, lo, block=1, grid_eff=0.7, max_size=32, mask=None):
    """Boxes (in the tags' index space, domain lower corner `lo`) covering the
    tagged cells of a dense (nz, ny, nx) array, within the region `mask`
    allows (same shape) when given."""
    dense = np.ascontiguousarray(dense)
    dense = dense.view(np.uint8) if dense.dtype == np.bool_ else dense.astype(np.uint8)
    nz, ny, nx = dense.shape
    lo_c = (ctypes.c_int * 3)(*lo)
    hi_c = (ctypes.c_int * 3)(lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
    nb = ctypes.c_int()
    up = ctypes.POINTER(ctypes.c_ubyte)
    mk = None
    if mask is not None:
        mask = np.ascontiguousarray(mask)
        mask = mask.view(np.uint8) if mask.dtype == np.bool_ else mask.astype(np.uint8)
        mk = mask.ctypes.data_as(up)
    cap = 4096
    while True:
        out, outp = box_buffer(cap)
        rc = lib().ocfd_cluster_tags(lo_c, hi_c, dense.ctypes.data_as(up), mk, block, grid_eff, max_size, cap, outp,
                                     ctypes.byref(nb))
        if rc != 0 and nb.value > cap:
            cap = nb.value          # more boxes than room: once more with the exact count
            continue
        check(rc, "ocfd_cluster_tags")
        break
    return boxes_from_buffer(out, nb.value)


def copy_level(dst, S_dst, src, S_src, stream=None):
    check(lib().ocfd_copy_level(dst.handle, _ptr(S_dst), src.handle, _ptr(S_src), _stream(stream)), "ocfd_copy_level")


def _grow(mask, n, periodic, shrink=False):
    """mask grown (or shrunk) by n cells, Chebyshev distance; periodic axes
    wrap, other sides see unset (grow) or set (shrink) cells.  n == 0: the
    mask itself as bool (no copy when it already is one)."""
    if n == 0:
        return np.asarray(mask, dtype=bool)
    a = np.ascontiguousarray(mask, dtype=np.uint8).copy()
    if n > 0:
        nz, ny, nx = a.shape
        per = (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic])
        check(lib().ocfd_mask_grow(a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), nx, ny, nz, n, per,
                                   int(shrink)), "ocfd_mask_grow")
    return a.view(bool)


def _dilate(mask, n, periodic):
    return _grow(mask, n, periodic)


def coverage(boxes, geom):
    """Dense (nz, ny, nx) coverage of a level's domain by its boxes."""
    lo = [geom.dom_lo[d] for d in range(3)]
    n = [geom.dom_hi[d] - lo[d] + 1 for d in range(3)]
    m = np.zeros((n[2], n[1], n[0]), dtype=bool)
    for blo, bhi in boxes:
        m[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1, blo[0] - lo[0]:bhi[0] - lo[0] + 1] = True
    return m


def proper_nesting(boxes, geom, n_proper):
    """Cells of a level whose n_proper-neighbourhood (inside the domain, periodic
    images included) is covered by the level: where the next finer level may go."""
    per = [geom.periodic[d] for d in range(3)]
    return _grow(coverage(boxes, geom), n_proper, per, shrink=True)


def projected_coverage(boxes_f2, geom_f, geom_c, n_proper):
    """Cells of level c (2 coarser than the boxes' level f+1) under the grids
    of level f+1 coarsened to level f, grown by n_proper there (periodic images
    included) and coarsened again: the tags that keep level f's new grids
    around the finer level's (Amr::grid_places, projection of new grids).
    Growing and coarsening distribute over a union of boxes, so this is box
    arithmetic and one coverage fill."""
    return coverage(projected_boxes(boxes_f2, geom_f, n_proper), geom_c)


def projected_boxes(boxes_f2, geom_f, n_proper):
    """The boxes of projected_coverage (level c index space, may overlap)."""
    n, lo, hi = box_corners(boxes_f2)
    dlo = (ctypes.c_int * 3)(*[geom_f.dom_lo[d] for d in range(3)])
    dhi = (ctypes.c_int * 3)(*[geom_f.dom_hi[d] for d in range(3)])
    per = (ctypes.c_int * 3)(*[int(bool(geom_f.periodic[d])) for d in range(3)])
    nb = ctypes.c_int()
    cap = max(64, 27 * n)
    out, outp = box_buffer(cap)
    check(lib().ocfd_project_boxes(n, lo, hi, dlo, dhi, per, int(n_proper), cap, outp, ctypes.byref(nb)),
          "ocfd_project_boxes")
    return boxes_from_buffer(out, nb.value)


def _spans_z(level):
    """Every box of the level spans its domain in z: then no coarse/fine face
    of the level faces z (a coarse column is wholly covered or not at all),
    and the register's face list need not be searched for one."""
    zlo, zhi = level.geom.dom_lo[2], level.geom.dom_hi[2]
    return all(lo[2] == zlo and hi[2] == zhi for lo, hi in level.boxes)


def coarsen_boxes(boxes):
    r = RATIO
    return [((a // r, b // r, c // r), (d // r, f // r, g // r)) for (a, b, c), (d, f, g) in boxes]


class FluxRegister:
    """The coarse/fine interface of one fine level (EBR's flux_reg), with the
    coarse-to-fine interpolation of that pair."""

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

    def fill_fine(self, S_fine, S_coarse, stream=None):
        """Fine ghosts no fine box owns, from the coarse state (ghosts filled)."""
        check(lib().ocfd_cf_fill_fine(self.handle, _ptr(S_fine), _ptr(S_coarse), _stream(stream)), "ocfd_cf_fill_fine")

    def interp_fine(self, S_fine, S_coarse, old_fine=None, stream=None):
        """Valid fine cells from the coarse state, except where old_fine covers them."""
        check(lib().ocfd_cf_interp_fine(self.handle, _ptr(S_fine), _ptr(S_coarse),
                                        old_fine.handle if old_fine is not None else None, _stream(stream)),
              "ocfd_cf_interp_fine")


class LevelState(ctypes.Structure):
    """struct ocfd_amr_level_state."""
    _fields_ = [("old", ctypes.c_void_p), ("cur", ctypes.c_void_p), ("s1", ctypes.c_void_p), ("s2", ctypes.c_void_p),
                ("tmp", ctypes.c_void_p), ("t_old", ctypes.c_double), ("t_new", ctypes.c_double)]


class Hierarchy:
    """A base level covering the domain and up to max_level finer levels of
    ratio 2, advanced with subcycling.

    coarse_boxes, geom: level 0.  kind: case initial data of ocfd_init_state
    (None: set level 0 with set_state and call start()).  fine_boxes: fixed
    grids -- a list per finer level, each in the index space of the level
    below (refined when built) -- instead of grids from tagging.  regrid_int 0
    keeps the grids.  blocking_factor and max_grid_size are in the finer
    level's cells (amr.* inputs); n_error_buf, n_proper, grid_eff as in AMReX.
    """

    def __init__(self, coarse_boxes, geom, parm, cfl, kind=None, scheme=2, flags=0, max_level=1,
                 refine_dengrad=0.05, regrid_int=2, blocking_factor=2, grid_eff=0.7, max_grid_size=64,
                 n_error_buf=0, n_proper=1, fine_boxes=None, seed=0, stop_time=-1.0, do_reflux=True, stream=None,
                 planar_fill=True):
        if scheme not in (0, 1, 2):
            raise ValueError("scheme must be 0 (Euler), 1 (RK2) or 2 (SSP-RK3)")
        if blocking_factor % RATIO or max_grid_size % blocking_factor:
            raise ValueError("blocking_factor must be a multiple of 2 and divide max_grid_size")
        if stream is None:
            import torch
            stream = torch.cuda.current_stream()
        self.parm, self.cfl, self.scheme, self.flags = parm, cfl, scheme, flags
        self.max_level = max_level
        self.thr, self.regrid_int = refine_dengrad, regrid_int
        self.block, self.grid_eff, self.max_c = blocking_factor // RATIO, grid_eff, max_grid_size // RATIO
        self.n_error_buf, self.n_proper = n_error_buf, n_proper
        self.stop_time, self.do_reflux, self.stream = stop_time, do_reflux, stream
        # PLANAR levels fill one z ghost layer and sweep their first z layer
        # (ocfd_level_set_planar); planar_fill=False keeps the four-layer fill
        self.planar_fill = planar_fill
        self.kind, self.seed = kind, seed
        if fine_boxes is not None and fine_boxes and not isinstance(fine_boxes[0][0][0], (tuple, list)):
            fine_boxes = [fine_boxes]          # one fixed fine level, boxes in level-0 index space
        self.fixed = fine_boxes
        self.geoms = [geom]
        for _ in range(max_level):
            self.geoms.append(refine_geom(self.geoms[-1]))
        n = max_level + 1
        self._retired = []
        self.levels = [None] * n
        self.cfs = [None] * n
        self.bufs = [None] * n
        self.st = (LevelState * n)()
        self.finest = 0
        self.time = 0.0
        self.dt_level = [0.0] * n
        self.level_steps = [0] * n
        self.count = [0] * n
        self.steps = 0
        self._set_level(0, Level(coarse_boxes, geom), None, 0.0)
        if kind is not None:
            self.levels[0].init_state(kind, self.S[0], parm, seed)
            self.start()

    # -- buffers ------------------------------------------------------------
    def _retire(self, *objs):
        """Keep replaced levels and registers alive until the next coarse step
        starts: freeing device memory waits for the whole device, which is idle
        then (the dt reductions have just synchronised) and busy during a
        regrid."""
        self._retired.extend(o for o in objs if o is not None)

    def _link(self, l, cf=None):
        if cf is None:
            cf = FluxRegister(self.levels[l - 1], self.levels[l])
        self._retire(self.cfs[l])
        if (self.flags & PLANAR) and cf.nfaces and not _spans_z(cf.fine) and (cf.faces()[:, 4] == 2).any():
            raise ValueError("planar stepping needs fine grids that span the domain in z")
        self.cfs[l] = cf

    def _set_level(self, l, level, S=None, t=0.0, cf=None, relink=True):
        """Install level l at time t (None removes it and everything finer);
        cf: its register with level l-1 when already built.  relink False:
        the next finer level is about to be rebuilt too, so it is not joined
        to the new level first."""
        if level is None:
            for m in range(l, self.max_level + 1):
                self._retire(self.levels[m], self.cfs[m])
                self.levels[m], self.bufs[m], self.cfs[m] = None, None, None
                self.st[m] = LevelState()
            self.finest = l - 1
            return
        self._retire(self.levels[l])
        self.levels[l] = level
        if (self.flags & PLANAR) and self.planar_fill:
            level.set_planar(True)
        # the old state, the stage states and the time-combined state have
        # their valid cells written and their ghost cells filled before any
        # read, so they start uninitialised (no fill kernels in a regrid)
        b = {k: level.new_mf(zero=False) for k in ("b", "s1", "s2", "tmp")}
        b["a"] = S if S is not None else level.new_mf()
        self.bufs[l] = b
        q = self.st[l]
        q.cur, q.old = b["a"].data_ptr(), b["b"].data_ptr()
        q.s1, q.s2, q.tmp = b["s1"].data_ptr(), b["s2"].data_ptr(), b["tmp"].data_ptr()
        q.t_old = q.t_new = t
        if l > 0:
            self._link(l, cf)
        # the next finer level now hangs off a new level; it is relinked, or
        # rebuilt by the caller when it no longer nests
        if l + 1 <= self.max_level and self.levels[l + 1] is not None:
            self._retire(self.cfs[l + 1])
            self.cfs[l + 1] = None
            if relink:
                try:
                    self._link(l + 1)
                except RuntimeError:
                    pass

    def _tensor(self, l, ptr):
        for t in self.bufs[l].values():
            if t.data_ptr() == ptr:
                return t
        raise RuntimeError("unknown state buffer")

    @property
    def S(self):
        """Current state of every existing level."""
        return [self._tensor(l, self.st[l].cur) for l in range(self.finest + 1)]

    @property
    def coarse(self):
        return self.levels[0]

    @property
    def fine(self):
        return self.levels[1] if self.finest >= 1 else None

    @property
    def gf(self):
        return self.geoms[1]

    def _arrays(self, n):
        """Handles of levels 0..n-1 and their registers."""
        lv = (ctypes.c_void_p * n)(*[self.levels[l].handle for l in range(n)])
        cf = (ctypes.c_void_p * n)(*[self.cfs[l].handle if l > 0 else None for l in range(n)])
        return n, lv, cf

    def fill(self, l, X=None, t=None):
        """Ghost cells of level l's state (default: its current one) at time t."""
        X = self.S[l] if X is None else X
        n, lv, cf = self._arrays(l + 1)
        check(lib().ocfd_amr_fill_level(n, lv, cf, self.st, l, _ptr(X), self.time if t is None else t,
                                        _stream(self.stream)), "ocfd_amr_fill_level")

    # -- grids ----------------------------------------------------------------
    def _tags(self, l, t):
        lev = self.levels[l]
        self.fill(l, None, t)
        tg = dense_tags(lev, tag_dengrad(lev, self.S[l], self.thr, self.stream).cpu().numpy()).astype(bool)
        g = self.geoms[l]
        return _dilate(tg, self.n_error_buf, [g.periodic[d] for d in range(3)])

    def _tag_flags(self, levc, t):
        """Device tag and nesting flags of level levc at time t (ghosts filled)."""
        self.fill(levc, None, t)
        return tag_dengrad_nest(self.levels[levc], self.S[levc], self.thr, self.n_proper if levc > 0 else 0,
                                self.stream)

    def _prefetch_flags(self, lbase, t):
        """The flags grid_places(lbase, t) reads, computed now and copied to
        pinned host memory without waiting: {levc: (level, host flags)} and an
        event that marks the copies' end."""
        import torch
        if self.n_error_buf != 0:
            return None
        out = {}
        with torch.cuda.stream(self.stream):
            for levc in range(min(self.finest, self.max_level - 1), lbase - 1, -1):
                dev = self._tag_flags(levc, t)
                host = torch.empty(dev.numel(), dtype=torch.uint8, pin_memory=True)
                host.copy_(dev, non_blocking=True)
                out[levc] = (self.levels[levc], host, dev)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        return out, ev

    def grid_places(self, lbase, t, pre=None):
        """New grids for levels lbase+1.. (Amr::grid_places): top-down from the
        finest level that may refine; each level's tags are widened to cover
        the next finer new grids plus n_proper cells, and kept inside the
        level's proper nesting region.  pre: the flags of _prefetch_flags(lbase, t)."""
        new = {}
        new_finest = lbase
        max_crse = min(self.finest, self.max_level - 1)
        for levc in range(max_crse, lbase - 1, -1):
            levf = levc + 1
            g = self.geoms[levc]
            if self.n_error_buf == 0:
                # tags and nesting region in the level's own layout (a level
                # covers a few percent of its domain), clustered from there
                lev = self.levels[levc]
                got = pre[0].get(levc) if pre is not None else None
                if got is not None and got[0] is lev:
                    pre[1].synchronize()
                    flags = got[1].numpy()
                else:
                    flags = self._tag_flags(levc, t).cpu().numpy()
                extra = []
                if levf + 1 in new and new[levf + 1]:
                    extra = projected_boxes(new[levf + 1], self.geoms[levf], self.n_proper)
                boxes = cluster_box_tags(lev.boxes, flags, extra, levc > 0, self.block, self.grid_eff, self.max_c)
                new[levf] = refine_boxes(boxes)
                if boxes:
                    new_finest = max(new_finest, levf)
                continue
            tags = self._tags(levc, t)
            if levf + 1 in new and new[levf + 1]:
                tags |= projected_coverage(new[levf + 1], self.geoms[levf], g, self.n_proper)
            pn = proper_nesting(self.levels[levc].boxes, g, self.n_proper) if levc > 0 else None
            if pn is not None:
                tags &= pn
            if not tags.any():
                new[levf] = []
                continue
            boxes = cluster_tags(tags, (g.dom_lo[0], g.dom_lo[1], g.dom_lo[2]), self.block, self.grid_eff, self.max_c,
                                 pn)
            new[levf] = refine_boxes(boxes)
            new_finest = max(new_finest, levf)
        return {l: b for l, b in new.items() if l <= new_finest}, new_finest

    def _fixed_grids(self):
        out, boxes = {}, None
        for l, fb in enumerate(self.fixed or [], start=1):
            if l > self.max_level:
                break
            out[l] = refine_boxes(fb)
        return out

    def _build(self, l, boxes, t, init_from_case, relink=True):
        """(Re)build level l on `boxes` at time t: case data, or the old level's
        cells where it had them and the interpolated coarser level elsewhere."""
        old_level = self.levels[l] if l <= self.finest else None
        old_S = self.S[l] if old_level is not None else None
        lev = Level(boxes, self.geoms[l])
        Sn = lev.new_mf()
        cf = FluxRegister(self.levels[l - 1], lev)
        if init_from_case and self.kind is not None:
            lev.init_state(self.kind, Sn, self.parm, self.seed)
        else:
            if old_level is not None:
                copy_level(lev, Sn, old_level, old_S, self.stream)
            self.fill(l - 1, None, t)
            cf.interp_fine(Sn, self.S[l - 1], old_level, self.stream)
        self._set_level(l, lev, Sn, t, cf, relink)
        self.finest = max(self.finest, l)

    def start(self):
        """Initial grids (Amr::bldFineLevels, then up to 4 top-down passes until
        the grids settle) and the average-down of EBR::post_init."""
        t = self.time
        if self.fixed is not None:
            for l, b in sorted(self._fixed_grids().items()):
                g = self.geoms[l - 1]
                inner = _dilate(coverage(coarsen_boxes(b), g), self.n_proper, [g.periodic[d] for d in range(3)])
                if (inner & ~coverage(self.levels[l - 1].boxes, g)).any():
                    raise ValueError(f"fixed grids of level {l} are not properly nested in level {l - 1} "
                                     f"(n_proper {self.n_proper}, periodic images included)")
                self._build(l, b, t, True)
        else:
            while self.finest < self.max_level:
                new, nf = self.grid_places(self.finest, t)
                if nf <= self.finest:
                    break
                self._build(nf, new[nf], t, True)
            if self.max_level > 0:
                for _ in range(4):
                    new, nf = self.grid_places(0, t)
                    if nf < self.finest:
                        break
                    same = True
                    for l in range(1, nf + 1):
                        if l > self.finest or sorted(new[l]) != sorted(self.levels[l].boxes):
                            same = False
                            self._build(l, new[l], t, True)
                    if same:
                        break
        self._check_links()
        for l in range(self.finest - 1, -1, -1):
            self.cfs[l + 1].avg_down(self.S[l + 1], self.S[l], NCONS, self.stream)

    def regrid(self, lbase, t, pre=None, then=None):
        """Amr::regrid(lbase): new grids for levels lbase+1..; levels that lose
        their grids are removed.  pre: flags from _prefetch_flags(lbase, t).
        then: called once level lbase+1 is in place, before the finer levels
        are built (the advance of level lbase, which reads levels up to
        lbase+1 only; the finer levels read the coarser ones at time t, which
        that advance keeps as its old state)."""
        if pre is None:
            # every level's tags queued at once: one wait for the first
            # clustering, the others' kernels run while it does
            pre = self._prefetch_flags(lbase, t)
        new, nf = self.grid_places(lbase, t, pre)
        for l in range(lbase + 1, nf + 1):
            if l == lbase + 2 and then is not None:
                then()
                then = None
            if new[l]:
                self._build(l, new[l], t, False, relink=not (l + 1 <= nf and new.get(l + 1)))
        if nf < self.finest:
            self._set_level(nf + 1, None)
        self._check_links()
        for l in range(lbase, self.max_level + 1):
            self.count[l] = 0
        if then is not None:
            then()

    def _check_links(self):
        for l in range(1, self.finest + 1):
            if self.cfs[l] is None:
                raise RuntimeError(f"level {l} is not nested in level {l - 1}")

    # -- time stepping ----------------------------------------------------------
    def _compute_dt(self):
        n = self.finest + 1
        lv = (ctypes.c_void_p * n)(*[self.levels[l].handle for l in range(n)])
        st = (ctypes.c_void_p * n)(*[self.S[l].data_ptr() for l in range(n)])
        out = (ctypes.c_double * n)()
        check(lib().ocfd_amr_estdt(n, lv, st, self.cfl, ctypes.byref(self.parm), out, _stream(self.stream)),
              "ocfd_amr_estdt")
        est = list(out)
        if self.steps > 0:
            prev = [self.dt_level[l] if self.dt_level[l] > 0 else self.dt_level[0] / RATIO ** l
                    for l in range(self.finest + 1)]
            est = [min(e, CHANGE_MAX * prev[l]) for l, e in enumerate(est)]
        dt0 = min(RATIO ** l * e for l, e in enumerate(est))
        eps = 0.001 * dt0
        if self.stop_time >= 0.0 and self.time + dt0 > self.stop_time - eps:
            dt0 = self.stop_time - self.time
        self.dt_level = [dt0 / RATIO ** l for l in range(self.max_level + 1)]

    def _advance(self, l, t, dt):
        n, lv, cf = self._arrays(min(l + 2, self.finest + 1))
        check(lib().ocfd_amr_advance_level(n, lv, cf, self.st, l, t, dt, self.scheme, ctypes.byref(self.parm),
                                           self.flags, int(self.do_reflux), _stream(self.stream)),
              "ocfd_amr_advance_level")

    def _time_step(self, l, t):
        # Amr::timeStep: regrid(i, t) at the first level i >= l whose count
        # reached regrid_int (it resets the counts of i and above, so no other
        # follows).  A regrid of a finer level i > l changes only levels above
        # i and their registers, while this level's advance reads levels up to
        # l+1 and the register of l+1, and the coarser states the regrid reads
        # at time t are the same before and after the advance (the advanced
        # level keeps its time-t state, ghosts filled, as its old state): so
        # its tags are computed and copied to the host first, the advance is
        # queued, and the clustering and the new levels are built on the host
        # while the device runs the advance.  Same results as regridding first.
        # A regrid of this level itself builds level l+1 first, queues the
        # advance, then builds the finer levels.
        pending = None
        advanced = False

        def advance():
            nonlocal advanced
            self._advance(l, t, self.dt_level[l])
            advanced = True

        if self.fixed is None and self.regrid_int > 0 and l < self.max_level:
            i = l
            while i <= min(self.finest, self.max_level - 1):
                if self.count[i] >= self.regrid_int:
                    if i == l:
                        self.regrid(i, t, then=advance)
                    else:
                        pending = (i, self._prefetch_flags(i, t))
                    break
                i += 1
        if not advanced:
            self._advance(l, t, self.dt_level[l])
        if pending is not None:
            self.regrid(pending[0], t, pending[1])
        self.level_steps[l] += 1
        self.count[l] += 1
        if l < self.finest:
            for it in range(RATIO):
                self._time_step(l + 1, t + it * self.dt_level[l + 1])
            if self.do_reflux:
                self.cfs[l + 1].reflux(self.S[l], self.stream)
            self.cfs[l + 1].avg_down(self.S[l + 1], self.S[l], NCONS, self.stream)

    def step(self):
        """One coarse time step (Amr::coarseTimeStep)."""
        self._compute_dt()
        self._retired.clear()
        t = self.time
        self._time_step(0, t)
        self.time = t + self.dt_level[0]
        self.steps += 1
        return self.dt_level[0]

    def run(self, nsteps=-1):
        n = 0
        while (nsteps < 0 or n < nsteps) and (self.stop_time < 0.0 or self.time < self.stop_time):
            self.step()
            n += 1
        return n

    # -- checkpoint / restart -----------------------------------------------------
    def checkpoint(self, path):
        """Amr::checkPoint (AmrLevel::checkPoint per level, Source/IO.cpp:27-30):
        a text Header with the hierarchy's clock and grids, and each level's
        current state (valid cells) as Level_<l>/SD_0_New in VisMF form.  The
        old states are not kept: every advance overwrites them before use."""
        os.makedirs(path, exist_ok=True)
        with open(os.path.join(path, "Header"), "w") as f:
            f.write("ocfd-checkpoint 1\n")
            f.write("max_level %d\nfinest_level %d\n" % (self.max_level, self.finest))
            f.write("time %r\nsteps %d\n" % (self.time, self.steps))
            f.write("dt_level %s\n" % " ".join(repr(float(x)) for x in self.dt_level))
            f.write("level_steps %s\n" % " ".join("%d" % x for x in self.level_steps))
            f.write("level_count %s\n" % " ".join("%d" % x for x in self.count))
            for l in range(self.finest + 1):
                f.write("level %d %d\n" % (l, len(self.levels[l].boxes)))
                for lo, hi in self.levels[l].boxes:
                    f.write("%d %d %d %d %d %d\n" % (*lo, *hi))
        for l in range(self.finest + 1):
            ldir = os.path.join(path, "Level_%d" % l)
            check(lib().ocfd_checkpoint_write_level(ldir.encode(), b"SD_0_New", self.levels[l].handle,
                                                    _ptr(self.S[l]), NCONS, _stream(self.stream)),
                  "ocfd_checkpoint_write_level")

    @classmethod
    def restart(cls, path, geom, parm, cfl, **kw):
        """Amr::restart: a hierarchy with the grids, clock and states of the
        checkpoint at `path` (geom: level 0's geometry; the remaining keywords
        as for the constructor, max_level defaulting to the checkpoint's)."""
        with open(os.path.join(path, "Header")) as f:
            lines = f.read().split("\n")
        if lines[0] != "ocfd-checkpoint 1":
            raise ValueError(f"{path} is not a checkpoint")
        hdr, boxes, i = {}, [], 1
        while i < len(lines) and lines[i]:
            key, *vals = lines[i].split()
            i += 1
            if key == "level":
                nb = int(vals[1])
                boxes.append([(tuple(int(x) for x in lines[i + b].split()[:3]),
                               tuple(int(x) for x in lines[i + b].split()[3:])) for b in range(nb)])
                i += nb
            else:
                hdr[key] = vals
        finest = int(hdr["finest_level"][0])
        kw.setdefault("max_level", int(hdr["max_level"][0]))
        if kw["max_level"] < finest:
            raise ValueError(f"max_level {kw['max_level']} is below the checkpoint's finest level {finest}")
        kw.pop("kind", None)
        h = cls(boxes[0], geom, parm, cfl, None, **kw)
        t = float(hdr["time"][0])
        for l in range(finest + 1):
            lev = h.levels[0] if l == 0 else Level(boxes[l], h.geoms[l])
            S = h.S[0] if l == 0 else lev.new_mf()
            check(lib().ocfd_checkpoint_read_level(os.path.join(path, "Level_%d" % l).encode(), b"SD_0_New",
                                                   lev.handle, _ptr(S), NCONS, _stream(h.stream)),
                  "ocfd_checkpoint_read_level")
            if l == 0:
                h.st[0].t_old = h.st[0].t_new = t
            else:
                h._set_level(l, lev, S, t)
                h.finest = l
        h._check_links()
        h.time, h.steps = t, int(hdr["steps"][0])
        n = h.max_level + 1

        def ext(vals, conv, fill):
            v = [conv(x) for x in vals][:n]
            return v + [fill] * (n - len(v))
        h.dt_level = ext(hdr["dt_level"], float, 0.0)
        h.level_steps = ext(hdr["level_steps"], int, 0)
        h.count = ext(hdr["level_count"], int, 0)
        return h

    def cells(self):
        """Valid cells per level."""
        out = []
        for l in range(self.finest + 1):
            out.append(sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
                           for lo, hi in self.levels[l].boxes))
        return out


def TwoLevel(coarse_boxes, geom, parm, cfl, kind=None, **kw):
    """A hierarchy of one base level and at most one finer level."""
    return Hierarchy(coarse_boxes, geom, parm, cfl, kind, max_level=1, **kw)
