"""CPU restatement of the two-level AMR machinery -- test infrastructure only.

Only tests/ use this module, as the checker of the device AMR path.  It
restates, in numpy over the C oracle's level fill and box fluxes:

* tag_dengrad, Source/Tagging.H:8-36;
* fine-grid generation from tags (Berger-Rigoutsos signature clustering with
  AMReX ClusterList's cut rules, blocking factor, joining of boxes that share
  a face, max grid size);
* AMReX's cell-conservative linear interpolation (MC-limited central slopes,
  third-order one-sided slopes at external-Dirichlet domain faces, one min/max
  limiter per coarse cell) for fine ghost cells and new fine regions: the
  interpolater of State_Type, Source/Setup.cpp:196;
* average-down, EBR::avgDown, Source/EBR.cpp:535-559 (unit volume fractions);
* the flux register (stage-weighted coarse and fine interface fluxes,
  Source/Advance.cpp:344-360) and reflux, Source/EBR.cpp:322-330;
* the subcycled two-level time step of Amr::coarseTimeStep with
  EBR::computeInitialDt / computeNewDt (Source/EBR.cpp:189-306).

AMReX itself is not available, so its interpolation, clustering and register
steps are restated from its published algorithms; their exact operation order
is this restatement's own (see the parity notes of the tests).
"""
import numpy as np

import oracle as orc

NC = orc.NCONS
NG = orc.NGROW
RATIO = 2
CHANGE_MAX = 1.1

# BCType per physical BC code (Source/Setup.cpp:21-34):
# 0 interior, 1 ext_dir, 2 foextrap, 3 reflect_even, -1 reflect_odd
INT, EXT, FOEXT, EVEN, ODD = 0, 3, 2, 1, -1
SCALAR_BC = [INT, EXT, FOEXT, EVEN, EVEN, EVEN]
NORM_BC = [INT, EXT, FOEXT, ODD, ODD, ODD]
TANG_BC = [INT, EXT, FOEXT, EVEN, EVEN, ODD]

STAGES = {
    0: [(0.0, 1.0)],
    1: [(0.0, 0.5), (1.0, 0.5)],
    2: [(0.0, 1.0 / 6.0), (1.0, 1.0 / 6.0), (0.5, 2.0 / 3.0)],
}


def comp_bc(g, n, d, hi):
    phys = g.hi_bc[d] if hi else g.lo_bc[d]
    if n in (0, 4):
        return SCALAR_BC[phys]
    return NORM_BC[phys] if n - 1 == d else TANG_BC[phys]


def dom(g):
    return [g.dom_lo[d] for d in range(3)], [g.dom_hi[d] for d in range(3)]


def refine_ogeom(g):
    f = orc.OGeom()
    import ctypes
    ctypes.memmove(ctypes.byref(f), ctypes.byref(g), ctypes.sizeof(orc.OGeom))
    for d in range(3):
        f.dom_lo[d] = RATIO * g.dom_lo[d]
        f.dom_hi[d] = RATIO * g.dom_hi[d] + RATIO - 1
        f.dx[d] = g.dx[d] / RATIO
    return f


def refine_boxes(boxes):
    return [(tuple(RATIO * v for v in lo), tuple(RATIO * v + RATIO - 1 for v in hi)) for lo, hi in boxes]


# ---------------------------------------------------------------------------
# dense views of a level
# ---------------------------------------------------------------------------
def dense_grown(boxes, fabs, g, ncomp=NC):
    """(ncomp, nz+8, ny+8, nx+8) array over the domain grown by the ghost width:
    the grown FABs (ghosts filled), then every valid region on top."""
    lo, hi = dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.full((ncomp, n[2] + 2 * NG, n[1] + 2 * NG, n[0] + 2 * NG), np.nan)
    for (blo, bhi), f in zip(boxes, fabs):
        glo = [blo[d] - NG for d in range(3)]
        ghi = [bhi[d] + NG for d in range(3)]
        a0 = [max(glo[d], lo[d] - NG) for d in range(3)]
        a1 = [min(ghi[d], hi[d] + NG) for d in range(3)]
        src = f.a[:, a0[2] - glo[2]:a1[2] - glo[2] + 1, a0[1] - glo[1]:a1[1] - glo[1] + 1,
                  a0[0] - glo[0]:a1[0] - glo[0] + 1]
        out[:, a0[2] - lo[2] + NG:a1[2] - lo[2] + NG + 1, a0[1] - lo[1] + NG:a1[1] - lo[1] + NG + 1,
            a0[0] - lo[0] + NG:a1[0] - lo[0] + NG + 1] = src
    for (blo, bhi), f in zip(boxes, fabs):
        out[:, blo[2] - lo[2] + NG:bhi[2] - lo[2] + NG + 1, blo[1] - lo[1] + NG:bhi[1] - lo[1] + NG + 1,
            blo[0] - lo[0] + NG:bhi[0] - lo[0] + NG + 1] = f.valid(blo, bhi)
    return out


def owner_mask(boxes, g):
    lo, hi = dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    m = np.zeros((n[2], n[1], n[0]), dtype=bool)
    for blo, bhi in boxes:
        m[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1, blo[0] - lo[0]:bhi[0] - lo[0] + 1] = True
    return m


# ---------------------------------------------------------------------------
# tagging and clustering
# ---------------------------------------------------------------------------
def tag_dengrad(boxes, fabs, g, thr):
    """Dense (nz, ny, nx) tags of a level whose ghosts are filled."""
    lo, hi = dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((n[2], n[1], n[0]), dtype=np.uint8)
    for (blo, bhi), f in zip(boxes, fabs):
        r = f.a[0]
        c = r[NG:-NG, NG:-NG, NG:-NG]
        ax = np.maximum(np.abs(r[NG:-NG, NG:-NG, NG + 1:r.shape[2] - NG + 1] - c),
                        np.abs(c - r[NG:-NG, NG:-NG, NG - 1:-NG - 1]))
        ay = np.maximum(np.abs(r[NG:-NG, NG + 1:r.shape[1] - NG + 1, NG:-NG] - c),
                        np.abs(c - r[NG:-NG, NG - 1:-NG - 1, NG:-NG]))
        az = np.maximum(np.abs(r[NG + 1:r.shape[0] - NG + 1, NG:-NG, NG:-NG] - c),
                        np.abs(c - r[NG - 1:-NG - 1, NG:-NG, NG:-NG]))
        t = (np.maximum(ax, np.maximum(ay, az)) >= thr).astype(np.uint8)
        out[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1, blo[0] - lo[0]:bhi[0] - lo[0] + 1] = t
    return out


def _find_cut(hist):
    """(cut offset, kind) with kind 0 hole, 1 inflection, 2 bisection, 3 none."""
    n = len(hist)
    if n <= 1:
        return 0, 3
    mid = n // 2
    cut, kind = -1, 3
    for i in range(n):
        if hist[i] == 0:
            kind = 0
            if abs(cut - mid) > abs(i - mid):
                cut = i
                if i > mid:
                    break
    if kind == 0:
        return cut, 0
    d2 = [0] * n
    for i in range(1, n - 1):
        d2[i] = hist[i + 1] - 2 * hist[i] + hist[i - 1]
    locmax = -1
    for i in range(2, n - 2):
        a, b = d2[i - 1], d2[i]
        dif = abs(a - b)
        if a * b < 0 and dif >= locmax:
            if dif > locmax:
                kind, cut, locmax = 1, i, dif
            elif abs(i - mid) < abs(cut - mid):
                cut = i
    if locmax <= 2:
        cut, kind = mid, 2
    return cut, kind


def _bbox(pts):
    return pts.min(0), pts.max(0)


def merge_boxes(boxes):
    """Join boxes that share their extent in two directions and touch in the
    third until no pair joins (sorted by lower corner, z major)."""
    bx = [(list(a), list(b)) for a, b in boxes]
    changed = True
    while changed:
        changed = False
        bx.sort(key=lambda q: (q[0][2], q[0][1], q[0][0]))
        i = 0
        while i < len(bx):
            a = bx[i]
            j = i + 1
            while j < len(bx):
                b = bx[j]
                same = [a[0][d] == b[0][d] and a[1][d] == b[1][d] for d in range(3)]
                if sum(same) == 2:
                    d = same.index(False)
                    if b[0][d] == a[1][d] + 1 or a[0][d] == b[1][d] + 1:
                        a[0][d] = min(a[0][d], b[0][d])
                        a[1][d] = max(a[1][d], b[1][d])
                        del bx[j]
                        changed = True
                        continue
                j += 1
            i += 1
    return bx


def _br(pts, grid_eff):
    """Berger-Rigoutsos clusters of integer points (n, 3) at efficiency grid_eff."""
    clusters = [pts]
    q = 0
    while q < len(clusters):
        p = clusters[q]
        blo, bhi = _bbox(p)
        vol = int(np.prod(bhi - blo + 1))
        if len(p) / vol >= grid_eff:
            q += 1
            continue
        cuts, kinds = [], []
        for d in range(3):
            hist = np.bincount(p[:, d] - blo[d], minlength=bhi[d] - blo[d] + 1).tolist()
            c, k = _find_cut(hist)
            cuts.append(blo[d] + c)
            kinds.append(k)
        best = min(kinds)
        dirn, minlen = -1, -1
        for d in range(3):
            if kinds[d] != best:
                continue
            m = min(cuts[d] - blo[d], bhi[d] - cuts[d])
            if m >= minlen:
                dirn, minlen = d, m
        low = p[p[:, dirn] < cuts[dirn]]
        high = p[p[:, dirn] >= cuts[dirn]]
        clusters[q] = low
        clusters.append(high)
    return clusters


def cluster(dense, lo, block=1, grid_eff=0.7, max_size=32, mask=None):
    """Boxes covering the tagged cells of dense (nz, ny, nx) (index origin lo),
    within the region `mask` allows when given."""
    nz, ny, nx = dense.shape
    b = block
    blk = dense.reshape(nz // b, b, ny // b, b, nx // b, b).max(axis=(1, 3, 5))
    kk, jj, ii = np.nonzero(blk)
    if len(ii) == 0:
        return []
    pts = np.stack([ii + lo[0] // b, jj + lo[1] // b, kk + lo[2] // b], axis=1)
    clusters = _br(pts, grid_eff)
    if mask is not None:
        bm = mask.astype(bool).reshape(nz // b, b, ny // b, b, nx // b, b).all(axis=(1, 3, 5))
        kept = []
        for p in clusters:
            blo, bhi = _bbox(p)
            sub = bm[blo[2] - lo[2] // b:bhi[2] - lo[2] // b + 1, blo[1] - lo[1] // b:bhi[1] - lo[1] // b + 1,
                     blo[0] - lo[0] // b:bhi[0] - lo[0] // b + 1]
            if sub.all():
                kept.append(p)
                continue
            zk, zj, zi = np.nonzero(sub)
            if len(zi) == 0:
                continue
            inside = np.stack([zi + blo[0], zj + blo[1], zk + blo[2]], axis=1)
            kept += _br(inside, 1.0)
        clusters = kept
    boxes = []
    for p in clusters:
        blo, bhi = _bbox(p)
        boxes.append(([int(v) * b for v in blo], [(int(v) + 1) * b - 1 for v in bhi]))
    boxes = merge_boxes(boxes)
    for d in range(3):
        out = []
        for blo, bhi in boxes:
            length = bhi[d] - blo[d] + 1
            if length > max_size:
                ratio, bs, nlen = 1, max_size, length
                while bs % 2 == 0 and nlen % 2 == 0:
                    ratio, bs, nlen = ratio * 2, bs // 2, nlen // 2
                numblk = nlen // bs + (1 if nlen % bs else 0)
                sz, extra = nlen // numblk, nlen % numblk
                hi_d = bhi[d]
                for k in range(numblk - 1):
                    ks = (sz + 1 if k < extra else sz) * ratio
                    tlo, thi = list(blo), list(bhi)
                    tlo[d], thi[d] = hi_d - ks + 1, hi_d
                    out.append((tlo, thi))
                    hi_d -= ks
                bhi = list(bhi)
                bhi[d] = hi_d
            out.append((blo, bhi))
        boxes = out
    return [(tuple(lo_), tuple(hi_)) for lo_, hi_ in boxes]


# ---------------------------------------------------------------------------
# cell-conservative linear interpolation
# ---------------------------------------------------------------------------
def interp_dense(C, g):
    """Fine values over the whole fine domain from the dense grown coarse array
    C (ncomp, nz+8, ny+8, nx+8): returns (ncomp, 2nz, 2ny, 2nx)."""
    ncomp = C.shape[0]
    lo, hi = dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = np.zeros((ncomp, 2 * n[2], 2 * n[1], 2 * n[0]))

    def sh(u, d, o):
        # u shifted by o cells along direction d, restricted to the domain
        sl = [slice(NG, NG + n[2]), slice(NG, NG + n[1]), slice(NG, NG + n[0])]
        ax = 2 - d
        sl[ax] = slice(NG + o, NG + o + n[d])
        return u[tuple(sl)]

    for c in range(ncomp):
        u = C[c]
        u0 = sh(u, 0, 0)
        sl = []
        for d in range(3):
            um, up = sh(u, d, -1), sh(u, d, 1)
            dc = 0.5 * (up - um)
            if not g.periodic[d]:
                idx = [slice(None)] * 3
                if comp_bc(g, c, d, 0) == EXT:
                    idx[2 - d] = slice(0, 1)
                    t = tuple(idx)
                    dc[t] = (-(16.0 / 15.0) * um[t] + 0.5 * u0[t] + (2.0 / 3.0) * up[t] - 0.1 * sh(u, d, 2)[t])
                if comp_bc(g, c, d, 1) == EXT:
                    idx[2 - d] = slice(n[d] - 1, n[d])
                    t = tuple(idx)
                    dc[t] = ((16.0 / 15.0) * up[t] - 0.5 * u0[t] - (2.0 / 3.0) * um[t] + 0.1 * sh(u, d, -2)[t])
            df, db = 2.0 * (up - u0), 2.0 * (u0 - um)
            s = np.where(df * db >= 0.0, np.minimum(np.abs(df), np.abs(db)), 0.0)
            sl.append(np.copysign(1.0, dc) * np.minimum(s, np.abs(dc)))
        umax, umin = u0.copy(), u0.copy()
        for kk in (-1, 0, 1):
            for jj in (-1, 0, 1):
                for ii in (-1, 0, 1):
                    w = u[NG + kk:NG + kk + n[2], NG + jj:NG + jj + n[1], NG + ii:NG + ii + n[0]]
                    umax = np.maximum(umax, w)
                    umin = np.minimum(umin, w)
        nz_ = (sl[0] != 0.0) | (sl[1] != 0.0) | (sl[2] != 0.0)
        dumax = np.abs(sl[0]) * 0.25 + np.abs(sl[1]) * 0.25 + np.abs(sl[2]) * 0.25
        alpha = np.ones_like(u0)
        with np.errstate(divide="ignore", invalid="ignore"):
            m1 = nz_ & (dumax * alpha > umax - u0)
            alpha = np.where(m1, (umax - u0) / dumax, alpha)
            m2 = nz_ & (dumax * alpha > u0 - umin)
            alpha = np.where(m2, (u0 - umin) / dumax, alpha)
        for oz in (0, 1):
            for oy in (0, 1):
                for ox in (0, 1):
                    xo, yo, zo = (0.25 if ox else -0.25), (0.25 if oy else -0.25), (0.25 if oz else -0.25)
                    out[c, oz::2, oy::2, ox::2] = u0 + xo * (sl[0] * alpha) + yo * (sl[1] * alpha) + zo * (sl[2] * alpha)
    return out


def fill_from_coarse(fboxes, ffabs, gf, Fint):
    """Fine ghost cells whose source no fine box owns take the interpolated
    coarse value at the source (after the level's own fill)."""
    lo, hi = dom(gf)
    own = owner_mask(fboxes, gf)
    for (blo, bhi), f in zip(fboxes, ffabs):
        glo = [blo[d] - NG for d in range(3)]
        shape = f.a.shape[1:]
        K, J, I = np.meshgrid(*(np.arange(shape[a]) + glo[2 - a] for a in range(3)), indexing="ij")
        ghost = ~((I >= blo[0]) & (I <= bhi[0]) & (J >= blo[1]) & (J <= bhi[1]) & (K >= blo[2]) & (K <= bhi[2]))
        if gf.probtype == 1 and comp_bc(gf, 0, 1, 0) == EVEN:
            x = gf.prob_lo[0] + (I + 0.5) * gf.dx[0]
            ghost &= ~((J < lo[1]) & (x < 1.0 / 6.0))
        iv = [I[ghost], J[ghost], K[ghost]]
        out = [np.zeros_like(iv[0]), np.zeros_like(iv[0]), np.zeros_like(iv[0])]
        for d in range(3):
            v = iv[d]
            if gf.periodic[d]:
                v = lo[d] + np.mod(v - lo[d], hi[d] - lo[d] + 1)
            iv[d] = v
            out[d] = np.where(v < lo[d], -1, np.where(v > hi[d], 1, 0))
        kidx = [K[ghost] - glo[2], J[ghost] - glo[1], I[ghost] - glo[0]]
        for c in range(NC):
            src = [iv[0].copy(), iv[1].copy(), iv[2].copy()]
            sign = np.ones(len(src[0]))
            user = np.zeros(len(src[0]), dtype=bool)
            for d in range(3):
                for side, s in ((0, -1), (1, 1)):
                    m = out[d] == s
                    bc = comp_bc(gf, c, d, side)
                    edge = lo[d] if side == 0 else hi[d]
                    if bc == FOEXT:
                        src[d] = np.where(m, edge, src[d])
                    elif bc in (EVEN, ODD):
                        src[d] = np.where(m, 2 * edge - src[d] - 1 if side == 0 else 2 * edge - src[d] + 1, src[d])
                        if bc == ODD:
                            sign = np.where(m, -sign, sign)
                    elif bc == EXT:
                        user |= m
            inside = np.ones(len(src[0]), dtype=bool)
            for d in range(3):
                inside &= (src[d] >= lo[d]) & (src[d] <= hi[d])
            sel = ~user & inside
            s0 = [np.clip(src[d] - lo[d], 0, hi[d] - lo[d]) for d in range(3)]
            sel &= ~own[s0[2], s0[1], s0[0]]
            val = sign * Fint[c, s0[2], s0[1], s0[0]]
            f.a[c, kidx[0][sel], kidx[1][sel], kidx[2][sel]] = val[sel]


def interp_valid(fboxes, ffabs, gf, Fint, skip=None):
    """Valid fine cells from the interpolated coarse state, except where the
    dense mask `skip` is set."""
    lo, _ = dom(gf)
    for (blo, bhi), f in zip(fboxes, ffabs):
        sl = (slice(None), slice(blo[2] - lo[2], bhi[2] - lo[2] + 1), slice(blo[1] - lo[1], bhi[1] - lo[1] + 1),
              slice(blo[0] - lo[0], bhi[0] - lo[0] + 1))
        v = f.valid(blo, bhi)
        if skip is None:
            v[:] = Fint[sl]
        else:
            m = ~skip[sl[1:]]
            v[:, m] = Fint[sl][:, m]


def copy_level(nboxes, nfabs, oboxes, ofabs, g):
    lo, _ = dom(g)
    D = dense_grown(oboxes, ofabs, g)[:, NG:-NG, NG:-NG, NG:-NG]
    own = owner_mask(oboxes, g)
    for (blo, bhi), f in zip(nboxes, nfabs):
        sl = (slice(blo[2] - lo[2], bhi[2] - lo[2] + 1), slice(blo[1] - lo[1], bhi[1] - lo[1] + 1),
              slice(blo[0] - lo[0], bhi[0] - lo[0] + 1))
        m = own[sl]
        v = f.valid(blo, bhi)
        v[:, m] = D[(slice(None),) + sl][:, m]
    return own


def avg_down(fboxes, ffabs, cboxes, cfabs, gc):
    """Coarse cells under fine boxes: sum of the 8 fine cells in i-fastest
    order, divided by 8."""
    lo, _ = dom(gc)
    for (blo, bhi), f in zip(fboxes, ffabs):
        v = f.valid(blo, bhi)
        acc = np.zeros((v.shape[0], v.shape[1] // 2, v.shape[2] // 2, v.shape[3] // 2))
        for kk in (0, 1):
            for jj in (0, 1):
                for ii in (0, 1):
                    acc = acc + v[:, kk::2, jj::2, ii::2]
        acc = acc / 8.0
        clo = [blo[d] // 2 for d in range(3)]
        chi = [bhi[d] // 2 for d in range(3)]
        for (cl, ch), cf in zip(cboxes, cfabs):
            a0 = [max(clo[d], cl[d]) for d in range(3)]
            a1 = [min(chi[d], ch[d]) for d in range(3)]
            if any(a0[d] > a1[d] for d in range(3)):
                continue
            cf.valid(a0, a1)[:] = acc[:, a0[2] - clo[2]:a1[2] - clo[2] + 1, a0[1] - clo[1]:a1[1] - clo[1] + 1,
                                     a0[0] - clo[0]:a1[0] - clo[0] + 1]


# ---------------------------------------------------------------------------
# flux register
# ---------------------------------------------------------------------------
def dense_faces(boxes, faces, g):
    """Dense face-flux arrays per direction over the domain from per-box face
    arrays (5, nz(+1), ny(+1), nx(+1)); faces outside every box are 0."""
    lo, hi = dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    out = []
    for d in range(3):
        shp = [n[2], n[1], n[0]]
        shp[2 - d] += 1
        A = np.zeros((NC,) + tuple(shp))
        for (blo, bhi), fc in zip(boxes, faces):
            e = [bhi[q] + (1 if q == d else 0) for q in range(3)]
            A[:, blo[2] - lo[2]:e[2] - lo[2] + 1, blo[1] - lo[1]:e[1] - lo[1] + 1, blo[0] - lo[0]:e[0] - lo[0] + 1] = fc[d]
        out.append(A)
    return out


def interface(cboxes, fboxes, gc):
    """Coarse cells next to the fine level with their interface faces
    (d, side) in the fixed order x lo, x hi, y lo, y hi, z lo, z hi."""
    lo, hi = dom(gc)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    cov = np.zeros((n[2], n[1], n[0]), dtype=bool)
    for blo, bhi in fboxes:
        cov[blo[2] // 2 - lo[2]:bhi[2] // 2 - lo[2] + 1, blo[1] // 2 - lo[1]:bhi[1] // 2 - lo[1] + 1,
            blo[0] // 2 - lo[0]:bhi[0] // 2 - lo[0] + 1] = True

    def covered(iv):
        iv = list(iv)
        for d in range(3):
            if gc.periodic[d]:
                iv[d] = lo[d] + (iv[d] - lo[d]) % n[d]
            if iv[d] < lo[d] or iv[d] > hi[d]:
                return -1
        return 1 if cov[iv[2] - lo[2], iv[1] - lo[1], iv[0] - lo[0]] else 0

    cells = []
    for blo, bhi in cboxes:
        for K in range(blo[2], bhi[2] + 1):
            for J in range(blo[1], bhi[1] + 1):
                for I in range(blo[0], bhi[0] + 1):
                    if covered((I, J, K)) != 0:
                        continue
                    fl = []
                    for d in range(3):
                        for side in (0, 1):
                            nb = [I, J, K]
                            nb[d] += 1 if side else -1
                            if covered(nb) == 1:
                                fl.append((d, side))
                    if fl:
                        cells.append(((I, J, K), fl))
    return cells


def reflux(cells, cboxes, cfabs, gc, gf, RC, RF):
    """S_coarse += -s (mean of the 4 fine sub-face sums - coarse sum) * dxinv
    for every interface face of every listed coarse cell."""
    lo_f, hi_f = dom(gf)
    nf = [hi_f[d] - lo_f[d] + 1 for d in range(3)]
    lo, _ = dom(gc)
    for (I, J, K), fl in cells:
        b = next(m for m, (blo, bhi) in enumerate(cboxes)
                 if all(blo[d] <= (I, J, K)[d] <= bhi[d] for d in range(3)))
        U = cfabs[b].valid((I, J, K), (I, J, K))[:, 0, 0, 0].copy()
        for d, side in fl:
            cf = [I, J, K]
            cf[d] += side
            rc = RC[d][:, cf[2] - lo[2], cf[1] - lo[1], cf[0] - lo[0]]
            t1, t2 = (1 if d == 0 else 0), (1 if d == 2 else 2)
            subs = []
            for a2 in (0, 1):
                for a1 in (0, 1):
                    g = [2 * I, 2 * J, 2 * K]
                    g[d] = 2 * (I, J, K)[d] + 2 if side else 2 * (I, J, K)[d] - 1
                    g[t1] += a1
                    g[t2] += a2
                    if gf.periodic[d]:
                        g[d] = lo_f[d] + (g[d] - lo_f[d]) % nf[d]
                    face = list(g)
                    if not side:
                        face[d] += 1         # the fine cell's high face
                    subs.append(RF[d][:, face[2] - lo_f[2], face[1] - lo_f[1], face[0] - lo_f[0]])
            fine = (((subs[0] + subs[1]) + subs[2]) + subs[3]) * 0.25
            sgn = 1.0 if side else -1.0
            U = U + (-sgn) * (fine - rc) * (1.0 / gc.dx[d])
        cfabs[b].valid((I, J, K), (I, J, K))[:, 0, 0, 0] = U


# ---------------------------------------------------------------------------
# level stages and the two-level driver
# ---------------------------------------------------------------------------
def stage(boxes, cur, old, new, c_cur, c_old, c_dt, g, p, visc):
    """new(valid) = (c_cur*cur + c_old*old) + c_dt * dSdt(cur); returns the
    per-box face fluxes of cur."""
    dxi = [1.0 / g.dx[d] for d in range(3)]
    faces = []
    for b, (lo, hi) in enumerate(boxes):
        D, fc = orc.compute_dsdt_box(lo, hi, cur[b], p, dxi, visc, 0, faces=True)
        v = c_cur * cur[b].valid(lo, hi)
        if old is not None:
            v = v + c_old * old[b].valid(lo, hi)
        new[b].valid(lo, hi)[:] = v + c_dt * D
        faces.append(fc)
    return faces


def copy_fabs(fabs):
    out = []
    for f in fabs:
        n = orc.Fab(f.lo, f.hi)
        n.a[:] = f.a
        out.append(n)
    return out


def dilate(mask, n, periodic):
    """mask grown by n cells (Chebyshev neighbourhood); periodic axes wrap,
    other sides contribute nothing."""
    out = mask.copy()
    if n <= 0:
        return out
    nz, ny, nx = mask.shape
    for dz in range(-n, n + 1):
        for dy in range(-n, n + 1):
            for dx in range(-n, n + 1):
                sh = mask
                for ax, dd, d in ((0, dz, 2), (1, dy, 1), (2, dx, 0)):
                    if dd == 0:
                        continue
                    if periodic[d]:
                        sh = np.roll(sh, dd, axis=ax)
                    else:
                        z = np.zeros_like(sh)
                        src = [slice(None)] * 3
                        dst = [slice(None)] * 3
                        if dd > 0:
                            src[ax], dst[ax] = slice(0, sh.shape[ax] - dd), slice(dd, None)
                        else:
                            src[ax], dst[ax] = slice(-dd, None), slice(0, sh.shape[ax] + dd)
                        z[tuple(dst)] = sh[tuple(src)]
                        sh = z
                out |= sh
    return out


def coarsen_boxes(boxes):
    return [(tuple(v // 2 for v in lo), tuple(v // 2 for v in hi)) for lo, hi in boxes]


class Hierarchy:
    """Oracle of the device Hierarchy driver (same arguments, oracle geometry):
    Amr::coarseTimeStep / timeStep with subcycling, regrids at every level,
    grid_places with proper nesting and projection, FillPatch from coarser
    levels interpolated to the stage time, flux registers per interface,
    reflux and average-down."""

    def __init__(self, cboxes, gc, p, cfl, kind, scheme=2, visc=0, max_level=1, refine_dengrad=0.05, regrid_int=2,
                 blocking_factor=2, grid_eff=0.7, max_grid_size=64, n_error_buf=0, n_proper=1, fine_boxes=None,
                 seed=0, stop_time=-1.0, do_reflux=True):
        self.p, self.cfl, self.scheme, self.visc = p, cfl, scheme, visc
        self.max_level = max_level
        self.thr, self.regrid_int = refine_dengrad, regrid_int
        self.block, self.eff, self.max_c = blocking_factor // 2, grid_eff, max_grid_size // 2
        self.n_error_buf, self.n_proper = n_error_buf, n_proper
        if fine_boxes is not None and fine_boxes and not isinstance(fine_boxes[0][0][0], (tuple, list)):
            fine_boxes = [fine_boxes]
        self.fixed, self.kind, self.seed = fine_boxes, kind, seed
        self.stop_time, self.do_reflux = stop_time, do_reflux
        self.geoms = [gc]
        for _ in range(max_level):
            self.geoms.append(refine_ogeom(self.geoms[-1]))
        n = max_level + 1
        self.boxes = [list(cboxes)] + [None] * max_level
        self.S = [orc.new_level_fabs(self.boxes[0])] + [None] * max_level
        self.O = [None] * n
        self.t_old = [0.0] * n
        self.t_new = [0.0] * n
        self.RC = [None] * n
        self.RF = [None] * n
        self.finest = 0
        self.time, self.steps = 0.0, 0
        self.dt_level = [0.0] * n
        self.count = [0] * n
        orc.init_case(kind, self.boxes[0], self.S[0], gc, p, seed)
        self._start()

    # compatibility names of the two-level form
    @property
    def cboxes(self):
        return self.boxes[0]

    @property
    def fboxes(self):
        return self.boxes[1] if self.finest >= 1 else []

    @property
    def gc(self):
        return self.geoms[0]

    @property
    def gf(self):
        return self.geoms[1]

    # -- fills ------------------------------------------------------------------
    def fill(self, l, X, ts):
        orc.fill_ghost(self.boxes[l], X, self.geoms[l], ts)
        if l > 0:
            C = self.at(l - 1, ts)
            Fint = interp_dense(dense_grown(self.boxes[l - 1], C, self.geoms[l - 1]), self.geoms[l - 1])
            fill_from_coarse(self.boxes[l], X, self.geoms[l], Fint)

    def at(self, m, ts):
        t0, t1 = self.t_old[m], self.t_new[m]
        span = t1 - t0
        if abs(ts - t1) <= 1e-10 * span:
            self.fill(m, self.S[m], t1)
            return self.S[m]
        if abs(ts - t0) <= 1e-10 * span:
            return self.O[m]
        T = copy_fabs(self.O[m])
        for tf, of, nf in zip(T, self.O[m], self.S[m]):
            tf.a[:] = ((t1 - ts) / span) * of.a + ((ts - t0) / span) * nf.a
        self.fill(m, T, ts)
        return T

    # -- grids --------------------------------------------------------------------
    def _tags(self, l, t):
        self.fill(l, self.S[l], t)
        g = self.geoms[l]
        tg = tag_dengrad(self.boxes[l], self.S[l], g, self.thr).astype(bool)
        return dilate(tg, self.n_error_buf, [g.periodic[d] for d in range(3)])

    def grid_places(self, lbase, t):
        new, new_finest = {}, lbase
        for levc in range(min(self.finest, self.max_level - 1), lbase - 1, -1):
            levf = levc + 1
            g = self.geoms[levc]
            per = [g.periodic[d] for d in range(3)]
            tags = self._tags(levc, t)
            if new.get(levf + 1):
                gf = self.geoms[levf]
                proj = dilate(owner_mask(coarsen_boxes(new[levf + 1]), gf), self.n_proper,
                              [gf.periodic[d] for d in range(3)])
                nz, ny, nx = proj.shape
                tags |= proj.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).any(axis=(1, 3, 5))
            pn = ~dilate(~owner_mask(self.boxes[levc], g), self.n_proper, per) if levc > 0 else None
            if pn is not None:
                tags &= pn
            if not tags.any():
                new[levf] = []
                continue
            lo, _ = dom(g)
            new[levf] = refine_boxes(cluster(tags.astype(np.uint8), lo, self.block, self.eff, self.max_c, pn))
            new_finest = max(new_finest, levf)
        return {l: b for l, b in new.items() if l <= new_finest}, new_finest

    def _build(self, l, boxes, t, from_case):
        new = orc.new_level_fabs(boxes)
        if from_case:
            orc.init_case(self.kind, boxes, new, self.geoms[l], self.p, self.seed)
        else:
            skip = None
            if l <= self.finest and self.boxes[l]:
                skip = copy_level(boxes, new, self.boxes[l], self.S[l], self.geoms[l])
            self.fill(l - 1, self.S[l - 1], t)
            Fint = interp_dense(dense_grown(self.boxes[l - 1], self.S[l - 1], self.geoms[l - 1]), self.geoms[l - 1])
            interp_valid(boxes, new, self.geoms[l], Fint, skip)
        self.boxes[l], self.S[l], self.O[l] = list(boxes), new, None
        self.t_old[l] = self.t_new[l] = t
        self.finest = max(self.finest, l)

    def _drop(self, l):
        for m in range(l, self.max_level + 1):
            self.boxes[m], self.S[m], self.O[m] = None, None, None
        self.finest = l - 1

    def _start(self):
        t = self.time
        if self.fixed is not None:
            for l, fb in enumerate(self.fixed, start=1):
                if l > self.max_level:
                    break
                self._build(l, refine_boxes(fb), t, True)
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
                        if l > self.finest or sorted(new[l]) != sorted(self.boxes[l]):
                            same = False
                            self._build(l, new[l], t, True)
                    if same:
                        break
        for l in range(self.finest - 1, -1, -1):
            avg_down(self.boxes[l + 1], self.S[l + 1], self.boxes[l], self.S[l], self.geoms[l])

    def regrid(self, lbase, t):
        new, nf = self.grid_places(lbase, t)
        for l in range(lbase + 1, nf + 1):
            if new[l]:
                self._build(l, new[l], t, False)
        if nf < self.finest:
            self._drop(nf + 1)
        for l in range(lbase, self.max_level + 1):
            self.count[l] = 0

    # -- time stepping --------------------------------------------------------------
    def _compute_dt(self):
        est = []
        for l in range(self.finest + 1):
            g = self.geoms[l]
            est.append(orc.estdt(self.boxes[l], self.S[l], [g.dx[d] for d in range(3)], self.cfl, self.p))
        if self.steps > 0:
            prev = [self.dt_level[l] if self.dt_level[l] > 0 else self.dt_level[0] / 2 ** l
                    for l in range(self.finest + 1)]
            est = [min(e, CHANGE_MAX * prev[l]) for l, e in enumerate(est)]
        dt0 = min(2 ** l * e for l, e in enumerate(est))
        eps = 0.001 * dt0
        if self.stop_time >= 0.0 and self.time + dt0 > self.stop_time - eps:
            dt0 = self.stop_time - self.time
        self.dt_level = [dt0 / 2 ** l for l in range(self.max_level + 1)]

    def _advance(self, l, t, dt):
        g = self.geoms[l]
        self.O[l] = self.S[l]
        N = copy_fabs(self.O[l])
        self.t_old[l], self.t_new[l] = t, t + dt
        if l + 1 <= self.finest:
            self.RC[l + 1] = None
            self.RF[l + 1] = None
        regs = self._run_stages(self.boxes[l], self.O[l], g, t, dt, lambda X, ts: self.fill(l, X, ts), N)
        self.S[l] = N
        if not self.do_reflux:
            return
        for w, fc in regs:
            if l + 1 <= self.finest:
                D = dense_faces(self.boxes[l], fc, g)
                self.RC[l + 1] = [w * D[d] if self.RC[l + 1] is None else self.RC[l + 1][d] + w * D[d]
                                  for d in range(3)]
            if l > 0:
                D = dense_faces(self.boxes[l], fc, g)
                self.RF[l] = [w * D[d] if self.RF[l] is None else self.RF[l][d] + w * D[d] for d in range(3)]

    def _run_stages(self, boxes, A, g, t, dt, fill, out):
        p, v = self.p, self.visc
        B, C = orc.new_level_fabs(boxes), orc.new_level_fabs(boxes)
        st = STAGES[self.scheme]
        regs = []
        fill(A, t)
        if self.scheme == 0:
            regs.append((st[0][1] * dt, stage(boxes, A, None, out, 1.0, 0.0, dt, g, p, v)))
            return regs
        regs.append((st[0][1] * dt, stage(boxes, A, None, B, 1.0, 0.0, dt, g, p, v)))
        fill(B, t + dt)
        if self.scheme == 1:
            regs.append((st[1][1] * dt, stage(boxes, B, A, out, 0.5, 0.5, 0.5 * dt, g, p, v)))
            return regs
        regs.append((st[1][1] * dt, stage(boxes, B, A, C, 0.25, 0.75, 0.25 * dt, g, p, v)))
        fill(C, t + 0.5 * dt)
        regs.append((st[2][1] * dt, stage(boxes, C, A, out, 2.0 / 3.0, 1.0 / 3.0, (2.0 / 3.0) * dt, g, p, v)))
        return regs

    def _time_step(self, l, t):
        if self.fixed is None and self.regrid_int > 0 and l < self.max_level:
            i = l
            while i <= min(self.finest, self.max_level - 1):
                if self.count[i] >= self.regrid_int:
                    self.regrid(i, t)
                i += 1
        self._advance(l, t, self.dt_level[l])
        self.count[l] += 1
        if l < self.finest:
            for it in range(2):
                self._time_step(l + 1, t + it * self.dt_level[l + 1])
            if self.do_reflux:
                cells = interface(self.boxes[l], self.boxes[l + 1], self.geoms[l])
                reflux(cells, self.boxes[l], self.S[l], self.geoms[l], self.geoms[l + 1], self.RC[l + 1],
                       self.RF[l + 1])
            avg_down(self.boxes[l + 1], self.S[l + 1], self.boxes[l], self.S[l], self.geoms[l])

    def step(self):
        self._compute_dt()
        t = self.time
        self._time_step(0, t)
        self.time = t + self.dt_level[0]
        self.steps += 1
        return self.dt_level[0]


def TwoLevel(cboxes, gc, p, cfl, kind, **kw):
    return Hierarchy(cboxes, gc, p, cfl, kind, max_level=1, **kw)
