"""Pins of the AMR oracle (oracle/amr_oracle.py) and of the host-side grid
generation of the library.

AMReX is not vendored with the reference and no reference test covers its AMR
machinery, so the oracle is pinned by what the mathematics fixes:
* the cell-conservative linear interpolation reproduces linear data exactly,
  keeps every fine value inside the min/max of the parent's coarse
  neighbours, conserves the parent's value, and its one-sided boundary slope
  is the exact derivative of any quadratic whose ghost value is the face value;
* average-down of the interpolation gives the coarse field back;
* the grid generator covers every tag with disjoint boxes, stops at the grid
  efficiency, and finds the obvious decompositions (separate blobs, an L);
* reflux makes the two-level periodic advance conserve mass, momentum and
  energy to rounding (and without reflux it does not);
* a fine level covering the whole domain reproduces the uniform fine run bit
  for bit, which pins the subcycled stage sequence and time interpolation.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import amr_oracle as ao


def _og(orc, n, periodic=(1, 1, 1), lo_bc=(0, 0, 0), hi_bc=(0, 0, 0), L=1.0):
    return orc.geom(n, (0.0, 0.0, 0.0), tuple(L * n[d] / n[0] for d in range(3)), periodic=periodic, lo_bc=lo_bc,
                    hi_bc=hi_bc)


def _dense_from_func(g, fn, ncomp=5):
    lo, hi = ao.dom(g)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    K, J, I = np.meshgrid(*(np.arange(-4, n[a] + 4) for a in (2, 1, 0)), indexing="ij")
    return np.stack([fn(I, J, K, c) for c in range(ncomp)])


def test_interp_reproduces_linear_data(orc):
    g = _og(orc, (8, 6, 4), periodic=(0, 0, 0), lo_bc=(2, 2, 2), hi_bc=(2, 2, 2))
    coef = np.array([[1.0, 0.3, -0.2, 0.05], [2.0, -0.1, 0.4, 0.0], [0.5, 0.0, 0.0, 0.7], [3.0, 1.0, 1.0, 1.0],
                     [-1.0, 0.2, 0.1, -0.3]])
    C = _dense_from_func(g, lambda I, J, K, c: coef[c, 0] + coef[c, 1] * I + coef[c, 2] * J + coef[c, 3] * K)
    F = ao.interp_dense(C, g)
    k, j, i = np.meshgrid(*(np.arange(F.shape[a]) for a in (1, 2, 3)), indexing="ij")
    xc, yc, zc = (i + 0.5) / 2 - 0.5, (j + 0.5) / 2 - 0.5, (k + 0.5) / 2 - 0.5
    for c in range(5):
        want = coef[c, 0] + coef[c, 1] * xc + coef[c, 2] * yc + coef[c, 3] * zc
        assert np.abs(F[c] - want).max() < 1e-13


def test_interp_conserves_and_stays_bounded(orc):
    g = _og(orc, (6, 6, 6))
    rng = np.random.default_rng(5)
    C = rng.uniform(0.5, 2.0, (5, 14, 14, 14))
    F = ao.interp_dense(C, g)
    mean = sum(F[:, kk::2, jj::2, ii::2] for kk in (0, 1) for jj in (0, 1) for ii in (0, 1)) / 8.0
    assert np.abs(mean - C[:, 4:-4, 4:-4, 4:-4]).max() < 1e-14
    nb = np.lib.stride_tricks.sliding_window_view(C[:, 3:-3, 3:-3, 3:-3], (3, 3, 3), axis=(1, 2, 3))
    umax = nb.max(axis=(-1, -2, -3)).repeat(2, 1).repeat(2, 2).repeat(2, 3)
    umin = nb.min(axis=(-1, -2, -3)).repeat(2, 1).repeat(2, 2).repeat(2, 3)
    assert (F <= umax + 1e-14).all() and (F >= umin - 1e-14).all()


def test_interp_one_sided_slope_at_external_boundary(orc):
    """x-lo inflow (ext_dir): u = -x^2 + 4x with the ghost holding the face
    value; the boundary cell's children are u0 -+ u'(0)/4 exactly (MC and
    the min/max limiter inactive for this profile), where a central slope
    would give (u(1)-u(-1/2))/2 = 2.625 instead of 4."""
    g = _og(orc, (6, 4, 4), periodic=(0, 1, 1), lo_bc=(1, 0, 0), hi_bc=(2, 0, 0))

    def fn(I, J, K, c):
        x = np.where(I < 0, -0.5 * np.ones_like(I, dtype=float), I.astype(float))
        x = np.where(I < -1, -0.5, x)
        return -x * x + 4.0 * x + 10.0 * (c + 1)

    C = _dense_from_func(g, fn)
    F = ao.interp_dense(C, g)
    for c in range(5):
        u0 = 10.0 * (c + 1)
        assert np.allclose(F[c, :, :, 0], u0 - 1.0, rtol=0, atol=1e-13)
        assert np.allclose(F[c, :, :, 1], u0 + 1.0, rtol=0, atol=1e-13)


def test_avg_down_of_interp_is_identity(orc):
    g = _og(orc, (8, 8, 4))
    rng = np.random.default_rng(9)
    boxes = [((0, 0, 0), (7, 7, 3))]
    fabs = orc.new_level_fabs(boxes)
    fabs[0].valid(*boxes[0])[:] = rng.uniform(1, 2, (5, 4, 8, 8))
    orc.fill_ghost(boxes, fabs, g, 0.0)
    C = ao.dense_grown(boxes, fabs, g)
    F = ao.interp_dense(C, g)
    fboxes = [((0, 0, 0), (15, 15, 7))]
    ff = orc.new_level_fabs(fboxes)
    ff[0].valid(*fboxes[0])[:] = F
    cf = orc.new_level_fabs(boxes)
    ao.avg_down(fboxes, ff, boxes, cf, g)
    assert np.abs(cf[0].valid(*boxes[0]) - fabs[0].valid(*boxes[0])).max() < 1e-15


def _covers(boxes, dense, lo=(0, 0, 0)):
    cov = np.zeros_like(dense, dtype=int)
    for blo, bhi in boxes:
        cov[blo[2] - lo[2]:bhi[2] - lo[2] + 1, blo[1] - lo[1]:bhi[1] - lo[1] + 1, blo[0] - lo[0]:bhi[0] - lo[0] + 1] += 1
    return cov


@pytest.mark.parametrize("eff", [0.5, 0.7, 1.0])
def test_cluster_properties(eff):
    rng = np.random.default_rng(int(eff * 10))
    for trial in range(5):
        d = (rng.uniform(size=(4, 24, 40)) < 0.08).astype(np.uint8)
        d[:, 5:9, 10:30] = 1
        boxes = ao.cluster(d, (0, 0, 0), 1, eff, 10 ** 6)
        cov = _covers(boxes, d)
        assert cov.max() == 1                      # disjoint
        assert (cov[d == 1] == 1).all()            # every tag covered
        for blo, bhi in boxes:
            sub = d[blo[2]:bhi[2] + 1, blo[1]:bhi[1] + 1, blo[0]:bhi[0] + 1]
            assert sub.mean() >= eff - 1e-12
            # bounding boxes: every face plane holds a tag
            assert sub[0].any() and sub[-1].any() and sub[:, 0].any() and sub[:, -1].any()
            assert sub[:, :, 0].any() and sub[:, :, -1].any()


def test_cluster_known_answers():
    d = np.zeros((2, 16, 32), dtype=np.uint8)
    d[:, 2:6, 3:9] = 1
    d[:, 10:14, 20:30] = 1
    assert sorted(ao.cluster(d, (0, 0, 0), 1, 0.7, 64)) == [((3, 2, 0), (8, 5, 1)), ((20, 10, 0), (29, 13, 1))]
    L = np.zeros((1, 8, 8), dtype=np.uint8)
    L[0, 0:8, 0:2] = 1
    L[0, 0:2, 0:8] = 1
    boxes = ao.cluster(L, (0, 0, 0), 1, 1.0, 64)
    assert len(boxes) == 2 and _covers(boxes, L).sum() == L.sum()
    # blocking factor 2: boxes aligned to 2; max size chops a long box evenly
    d = np.zeros((2, 8, 64), dtype=np.uint8)
    d[:, 3, 5:61] = 1
    boxes = ao.cluster(d, (0, 0, 0), 2, 0.7, 16)
    for blo, bhi in boxes:
        assert all(v % 2 == 0 for v in blo) and all((v + 1) % 2 == 0 for v in bhi)
        assert all(bhi[q] - blo[q] + 1 <= 16 for q in range(3))
    assert sorted(b[0][0] for b in boxes) == [4, 18, 32, 46]


@pytest.mark.parametrize("block,eff,mx", [(1, 0.7, 32), (1, 1.0, 16), (2, 0.7, 8), (1, 0.5, 64)])
def test_library_clustering_matches_oracle(block, eff, mx):
    """ocfd_cluster_tags (host code of the HIP library) against the oracle."""
    from ocfd.amr import cluster_tags
    rng = np.random.default_rng(block * 100 + mx)
    for trial in range(6):
        shape = (2 * int(rng.integers(1, 4)), 2 * int(rng.integers(4, 20)), 2 * int(rng.integers(4, 40)))
        d = (rng.uniform(size=shape) < rng.uniform(0.02, 0.3)).astype(np.uint8)
        lo = (2 * int(rng.integers(0, 3)), 0, 2)
        want = sorted(ao.cluster(d, lo, block, eff, mx))
        got = sorted(cluster_tags(d, lo, block, eff, mx))
        assert got == want, (trial, shape)


def _kh_two_level(orc, n, fine, reflux=True, nsteps=3, scheme=2, visc=1):
    g = _og(orc, (n, n, n))
    boxes = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    T = ao.TwoLevel(boxes, g, orc.parm(), 0.4, 2, scheme=scheme, visc=visc, regrid_int=0, fine_boxes=fine,
                    seed=3, do_reflux=reflux)
    return T, g, boxes


def _totals(T):
    tot = np.zeros(5)
    cov = np.zeros((T.gc.dom_hi[2] + 1, T.gc.dom_hi[1] + 1, T.gc.dom_hi[0] + 1), dtype=bool)
    for blo, bhi in T.fboxes:
        cov[blo[2] // 2:bhi[2] // 2 + 1, blo[1] // 2:bhi[1] // 2 + 1, blo[0] // 2:bhi[0] // 2 + 1] = True
    for (lo, hi), f in zip(T.cboxes, T.S[0]):
        v = f.valid(lo, hi)
        m = ~cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        tot += 8.0 * v[:, m].sum(axis=1)
    for (lo, hi), f in zip(T.fboxes, T.S[1]):
        tot += f.valid(lo, hi).reshape(5, -1).sum(axis=1)
    return tot


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_reflux_conserves_on_a_periodic_domain(orc, scheme):
    fine = [((2, 1, 2), (5, 4, 4)), ((6, 1, 2), (7, 6, 5))]
    T, g, boxes = _kh_two_level(orc, 12, fine, scheme=scheme)
    t0 = _totals(T)
    for _ in range(3):
        T.step()
    t1 = _totals(T)
    scale = np.array([t0[0], np.abs(t0[1:4]).max() + t0[0], 1, 1, t0[4]])
    scale[2] = scale[3] = scale[1]
    assert (np.abs(t1 - t0) / scale < 2e-14).all(), (t1 - t0) / scale
    U, _, _ = _kh_two_level(orc, 12, fine, reflux=False, scheme=scheme)
    u0 = _totals(U)
    for _ in range(3):
        U.step()
    assert np.abs(_totals(U) - u0)[0] / scale[0] > 1e-10


def test_fully_refined_equals_uniform_fine(orc):
    n = 6
    T, g, boxes = _kh_two_level(orc, n, [((0, 0, 0), (n - 1, n - 1, n - 1))], nsteps=0)
    gf = ao.refine_ogeom(g)
    fb = ao.refine_boxes(boxes)
    A = orc.new_level_fabs(fb)
    B = orc.new_level_fabs(fb)
    C = orc.new_level_fabs(fb)
    orc.init_case(2, fb, A, gf, orc.parm(), 3)
    time = 0.0
    for _ in range(3):
        dt0 = T.step()
        for it in range(2):
            orc.flow_advance(fb, A, B, C, gf, orc.parm(), time + it * dt0 / 2, dt0 / 2, 2, 1)
        time += dt0
    for (lo, hi), f, a in zip(fb, T.S[1], A):
        assert np.array_equal(f.valid(lo, hi), a.valid(lo, hi))


def _totals_h(H):
    """Totals of a hierarchy: every level's cells not covered by the next finer
    one, in units of the finest cell volume."""
    tot = np.zeros(5)
    for l in range(H.finest + 1):
        g = H.geoms[l]
        n = [g.dom_hi[d] + 1 for d in range(3)]
        cov = np.zeros((n[2], n[1], n[0]), dtype=bool)
        if l < H.finest:
            for lo, hi in H.boxes[l + 1]:
                cov[lo[2] // 2:hi[2] // 2 + 1, lo[1] // 2:hi[1] // 2 + 1, lo[0] // 2:hi[0] // 2 + 1] = True
        w = 8.0 ** (H.finest - l)
        for (lo, hi), f in zip(H.boxes[l], H.S[l]):
            v = f.valid(lo, hi)
            m = ~cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            tot += w * v[:, m].sum(axis=1)
    return tot


def test_three_levels_conserve_on_a_periodic_domain(orc):
    """Viscous shear layer, 8^3 base + properly nested fixed levels 1 and 2
    (level 2 across a periodic face): three levels of subcycling with reflux
    at both interfaces conserve every total to rounding."""
    n = 8
    g = _og(orc, (n, n, n))
    fine = [[((0, 1, 1), (7, 6, 5))], [((0, 4, 4), (3, 9, 7)), ((14, 4, 4), (15, 9, 7)), ((6, 3, 3), (9, 7, 5))]]
    H = ao.Hierarchy([((0, 0, 0), (n - 1, n - 1, n - 1))], g, orc.parm(), 0.4, 2, visc=1, max_level=2,
                     regrid_int=0, fine_boxes=fine, seed=5)
    assert H.finest == 2
    t0 = _totals_h(H)
    for _ in range(2):
        H.step()
    t1 = _totals_h(H)
    scale = np.array([t0[0]] * 4 + [t0[4]])
    assert (np.abs(t1 - t0) / scale < 2e-14).all(), (t1 - t0) / scale


def test_three_levels_fully_refined_equal_uniform_finest(orc):
    n = 4
    g = _og(orc, (n, n, n))
    fine = [[((0, 0, 0), (n - 1, n - 1, n - 1))], [((0, 0, 0), (2 * n - 1, 2 * n - 1, 2 * n - 1))]]
    H = ao.Hierarchy([((0, 0, 0), (n - 1, n - 1, n - 1))], g, orc.parm(), 0.4, 2, visc=1, max_level=2,
                     regrid_int=0, fine_boxes=fine, seed=3)
    g2 = ao.refine_ogeom(ao.refine_ogeom(g))
    fb = [((0, 0, 0), (4 * n - 1, 4 * n - 1, 4 * n - 1))]
    A, B, C = (orc.new_level_fabs(fb) for _ in range(3))
    orc.init_case(2, fb, A, g2, orc.parm(), 3)
    time = 0.0
    for _ in range(2):
        dt0 = H.step()
        for it in range(4):
            orc.flow_advance(fb, A, B, C, g2, orc.parm(), time + it * dt0 / 4, dt0 / 4, 2, 1)
        time += dt0
    assert np.array_equal(H.S[2][0].valid(*fb[0]), A[0].valid(*fb[0]))


def test_three_level_dmr_grids_nest(orc):
    """Double Mach reflection with max_level 2: after regrids every level's
    grids, coarsened and grown by n_proper, lie inside the level below (away
    from the domain boundary), and every tag of each level's last regrid is
    refined."""
    from ocfd import cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    og = orc.OGeom()
    ctypes.memmove(ctypes.byref(og), ctypes.byref(g), ctypes.sizeof(og))
    H = ao.Hierarchy(boxes, og, orc.parm(), 0.5, kind, max_level=2, regrid_int=2, max_grid_size=32)
    for _ in range(5):
        H.step()
    assert H.finest == 2
    for l in range(1, H.finest + 1):
        gc = H.geoms[l - 1]
        per = [gc.periodic[d] for d in range(3)]
        cov_c = ao.owner_mask(H.boxes[l - 1], gc)
        fine_c = ao.owner_mask(ao.coarsen_boxes(H.boxes[l]), gc)
        assert not (fine_c & ~cov_c).any()
        if l >= 2:
            grown = ao.dilate(fine_c, 1, per)
            assert not (grown & ~cov_c).any()


@pytest.mark.parametrize("block", [1, 2])
def test_clustering_within_a_mask(block):
    """Clusters are cut back to the allowed region (the proper nesting domain):
    every box lies inside the mask, every tag inside the mask is covered, and the
    library gives the oracle's boxes."""
    from ocfd.amr import cluster_tags
    rng = np.random.default_rng(40 + block)
    for trial in range(6):
        shape = (2, 2 * int(rng.integers(6, 16)), 2 * int(rng.integers(8, 24)))
        mask = np.zeros(shape, dtype=np.uint8)
        for _ in range(3):
            y0, x0 = 2 * int(rng.integers(0, shape[1] // 2 - 2)), 2 * int(rng.integers(0, shape[2] // 2 - 2))
            mask[:, y0:y0 + 2 * int(rng.integers(2, 6)), x0:x0 + 2 * int(rng.integers(2, 8))] = 1
        tags = ((rng.uniform(size=shape) < 0.3) & mask.astype(bool)).astype(np.uint8)
        want = sorted(ao.cluster(tags, (0, 0, 0), block, 0.7, 64, mask))
        got = sorted(cluster_tags(tags, (0, 0, 0), block, 0.7, 64, mask))
        assert got == want, trial
        cov = _covers(got, tags)
        assert cov.max() <= 1 and (cov[tags == 1] == 1).all()
        assert not (cov.astype(bool) & ~mask.astype(bool)).any()


def _brute_grow(m, n, periodic):
    """Chebyshev growth by n, one cell at a time (pure Python loops)."""
    nz, ny, nx = m.shape
    out = np.zeros_like(m)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                hit = False
                for dk in range(-n, n + 1):
                    for dj in range(-n, n + 1):
                        for di in range(-n, n + 1):
                            q = [i + di, j + dj, k + dk]
                            ok = True
                            for d, size in ((0, nx), (1, ny), (2, nz)):
                                if 0 <= q[d] < size:
                                    continue
                                if periodic[d]:
                                    q[d] %= size
                                else:
                                    ok = False
                            if ok and m[q[2], q[1], q[0]]:
                                hit = True
                out[k, j, i] = hit
    return out


@pytest.mark.parametrize("shape,periodic,n", [((3, 4, 5), (0, 0, 0), 1), ((3, 4, 5), (1, 0, 1), 1),
                                              ((2, 5, 6), (1, 1, 0), 2)])
def test_oracle_dilation_matches_brute_force(shape, periodic, n):
    rng = np.random.default_rng(sum(shape) + n)
    m = rng.random(shape) < 0.15
    assert np.array_equal(ao.dilate(m, n, periodic), _brute_grow(m, n, periodic))


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_library_mask_grow_and_shrink_match_oracle(periodic, n):
    """ocfd_mask_grow against the oracle's 27-shift dilation: growth of random
    tags, and shrinking (the proper nesting domain) as the complement of the
    growth of the complement."""
    from ocfd import amr
    rng = np.random.default_rng(n * 7 + sum(periodic))
    for shape in ((4, 9, 11), (1, 6, 13), (7, 1, 5)):
        m = rng.random(shape) < 0.2
        assert np.array_equal(amr._grow(m, n, periodic), ao.dilate(m, n, periodic))
        c = rng.random(shape) < 0.8
        assert np.array_equal(amr._grow(c, n, periodic, shrink=True), ~ao.dilate(~c, n, periodic))


def test_projected_coverage_matches_dense_projection():
    """Grown-and-coarsened finer grids as box arithmetic equal the dense
    route: coverage at level f, Chebyshev growth (periodic images), any-reduce
    by the refinement ratio."""
    import ocfd
    from ocfd import amr
    rng = np.random.default_rng(3)
    for per in ((0, 0, 0), (1, 1, 0), (1, 1, 1)):
        gc = ocfd.Geom()
        for d in range(3):
            gc.dom_lo[d], gc.periodic[d] = 0, per[d]
        gc.dom_hi[0], gc.dom_hi[1], gc.dom_hi[2] = 15, 11, 7
        gf = amr.refine_geom(gc)
        gf2 = amr.refine_geom(gf)
        for n_proper in (1, 2):
            boxes = []
            for _ in range(6):
                lo = [int(rng.integers(0, (gf2.dom_hi[d] + 1) - 4)) for d in range(3)]
                hi = [lo[d] + int(rng.integers(1, 4)) for d in range(3)]
                boxes.append((tuple(lo), tuple(hi)))
            want = ao.dilate(amr.coverage(amr.coarsen_boxes(boxes), gf), n_proper, per)
            nz, ny, nx = want.shape
            want = want.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).any(axis=(1, 3, 5))
            assert np.array_equal(amr.projected_coverage(boxes, gf, gc, n_proper), want), (per, n_proper)


@pytest.mark.parametrize("masked", [False, True])
def test_library_clustering_matches_oracle_on_fronts(masked):
    """Oblique tagged fronts (the shape shock tags take), where clustering
    leaves many thin boxes and the box merge joins many pairs: library and
    oracle give the same boxes, with and without a nesting mask."""
    from ocfd.amr import cluster_tags
    rng = np.random.default_rng(21 + masked)
    for trial in range(8):
        shape = (2, 2 * int(rng.integers(16, 40)), 2 * int(rng.integers(24, 64)))
        d = np.zeros(shape, np.uint8)
        for _ in range(int(rng.integers(1, 4))):
            x0, sl, w = int(rng.integers(0, shape[2])), rng.uniform(-1.5, 1.5), int(rng.integers(1, 4))
            for j in range(shape[1]):
                x = int(x0 + sl * j)
                if 0 <= x < shape[2] - w:
                    d[:, j, x:x + w] = 1
        mask = None
        if masked:
            mask = np.ones(shape, np.uint8)
            mask[:, :, shape[2] // 2:] = 0
        for block, eff, mx in ((1, 0.7, 32), (2, 0.7, 8)):
            want = sorted(ao.cluster(d, (0, 0, 0), block, eff, mx, mask))
            got = sorted(cluster_tags(d, (0, 0, 0), block, eff, mx, mask))
            assert got == want, (trial, block)


@pytest.mark.parametrize("masked", [False, True])
def test_library_clustering_matches_oracle_on_a_large_front_set(masked):
    """Six oblique fronts over a 1024x256x2 domain, about 10 000 tagged
    points: enough for the library to refine the first clusters serially and
    the rest on its host threads. The boxes must still be the oracle's."""
    from ocfd.amr import cluster_tags
    rng = np.random.default_rng(5)
    shape = (2, 256, 1024)
    d = np.zeros(shape, np.uint8)
    for _ in range(6):
        x0, sl, w = int(rng.integers(0, shape[2])), rng.uniform(-1.5, 1.5), int(rng.integers(2, 6))
        for j in range(shape[1]):
            x = int(x0 + sl * j)
            if 0 <= x < shape[2] - w:
                d[:, j, x:x + w] = 1
    mask = None
    if masked:
        mask = np.ones(shape, np.uint8)
        mask[:, 96:160, 256:768] = 0
    assert int(d.sum()) > 8192
    want = sorted(ao.cluster(d, (0, 0, 0), 1, 0.7, 32, mask))
    got = sorted(cluster_tags(d, (0, 0, 0), 1, 0.7, 32, mask))
    assert len(want) > 256
    assert got == want


def _level_flags(boxes, dense_bits):
    """Flags in the level's 0-ghost layout (box order, i fastest) from a dense
    (nz, ny, nx) array over a domain starting at 0."""
    out = []
    for lo, hi in boxes:
        out.append(dense_bits[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].reshape(-1))
    return np.concatenate(out).astype(np.uint8)


@pytest.mark.parametrize("block", [1, 2])
def test_box_layout_clustering_equals_dense(block):
    """Clustering from tags in a level's own layout (with the proper nesting
    region in bit 1 and the projected finer grids as extra boxes) returns the
    boxes of the dense clustering of (tags | extra) & region within the
    region, and those of the oracle."""
    import ocfd
    from ocfd.amr import cluster_box_tags, cluster_tags, coverage, proper_nesting
    rng = np.random.default_rng(40 + block)
    for trial in range(10):
        nb = (int(rng.integers(3, 7)), int(rng.integers(2, 5)), 1)
        bs = 8
        n = (nb[0] * bs, nb[1] * bs, 4)
        periodic = (int(rng.integers(0, 2)), 0, 1)
        g = ocfd.make_geom(n, (0, 0, 0), (1.0, n[1] / n[0], n[2] / n[0]), periodic=periodic)
        full = trial % 3 == 0          # a base level covering the domain: no region
        boxes = []
        for bj in range(nb[1]):
            for bi in range(nb[0]):
                if full or rng.random() < 0.6:
                    boxes.append(((bi * bs, bj * bs, 0), (bi * bs + bs - 1, bj * bs + bs - 1, n[2] - 1)))
        if not boxes:
            continue
        shape = (n[2], n[1], n[0])
        tags = np.zeros(shape, np.uint8)
        for _ in range(int(rng.integers(1, 4))):
            x0, sl, w = int(rng.integers(0, n[0])), rng.uniform(-1.5, 1.5), int(rng.integers(1, 3))
            for j in range(n[1]):
                x = int(x0 + sl * j)
                if 0 <= x < n[0] - w:
                    tags[:, j, x:x + w] = 1
        tags |= (rng.random(shape) < 0.01).astype(np.uint8)
        cov = coverage(boxes, g).astype(np.uint8)
        tags &= cov                                   # a level tags its own cells only
        region = cov.copy() if full else proper_nesting(boxes, g, 1).astype(np.uint8)
        extra = []
        for _ in range(int(rng.integers(0, 3))):
            lo = [int(rng.integers(0, n[d] - 2)) for d in range(3)]
            hi = [min(n[d] - 1, lo[d] + int(rng.integers(0, 5))) for d in range(3)]
            extra.append((tuple(lo), tuple(hi)))
        dense = tags | coverage(extra, g).astype(np.uint8) if extra else tags.copy()
        dense &= region
        flags = _level_flags(boxes, tags | (region << 1))
        for eff, mx in ((0.7, 16), (0.9, 8)):
            if not dense.any():
                continue
            mask = None if full else region
            want = sorted(cluster_tags(dense, (0, 0, 0), block, eff, mx, mask))
            got = sorted(cluster_box_tags(boxes, flags, extra, not full, block, eff, mx))
            assert got == want, (trial, block, eff)
            assert got == sorted(ao.cluster(dense, (0, 0, 0), block, eff, mx, mask)), (trial, block, eff)


def test_clustering_on_host_threads_matches_serial():
    """Large tag sets are refined on the library's host threads: the boxes are
    those of the serial refinement (OCFD_HOST_THREADS=1, a process of its own)
    and of the oracle."""
    import subprocess
    import sys
    rng = np.random.default_rng(5)
    shape = (4, 256, 512)
    d = np.zeros(shape, np.uint8)
    for _ in range(6):
        x0, sl, w = int(rng.integers(0, shape[2])), rng.uniform(-1.5, 1.5), int(rng.integers(3, 9))
        for j in range(shape[1]):
            x = int(x0 + sl * j)
            if 0 <= x < shape[2] - w:
                d[:, j, x:x + w] = 1
    assert d.sum() > 4096 * 2
    code = ("import sys, numpy as np; sys.path.insert(0, %r); from ocfd.amr import cluster_tags; "
            "d = np.load(sys.argv[1]); print(sorted(cluster_tags(d, (0, 0, 0), 1, 0.7, 32)))"
            % str(Path(__file__).resolve().parent.parent / "opencfd-amr_amd"))
    import tempfile, os
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "d.npy")
        np.save(f, d)
        outs = []
        for nt in ("1", "8"):
            env = dict(os.environ, OCFD_HOST_THREADS=nt)
            r = subprocess.run([sys.executable, "-c", code, f], capture_output=True, text=True, env=env, timeout=120)
            assert r.returncode == 0, r.stderr
            outs.append(r.stdout)
    assert outs[0] == outs[1]
    want = str(sorted(ao.cluster(d, (0, 0, 0), 1, 0.7, 32))) + "\n"
    assert outs[0] == want
