"""Embedded boundary on the CPU: the cut-cell geometry generator (host code of
the library, replacing AMReX EB2::Build) pinned by exact areas of straight
boundaries, second-order convergence of curved ones and the closure of every
cut polygon; and the oracle's EB box path (Source/EB/AdvanceBox.cpp,
EBkernels.H, EBdiffusion.H) pinned by the Van Leer splitting's consistency,
a quiescent gas staying at rest next to the wall, conservation of mass and
energy on a periodic box around a body, and free-stream preservation away
from the body."""
import math

import numpy as np
import pytest

import ocfd
from ocfd import cases


def _geom(n, hi=(1.0, 1.0, None)):
    nx, ny, nz = n
    dx = hi[0] / nx
    return ocfd.make_geom(n, (0.0, 0.0, 0.0), (hi[0], hi[1], nz * dx), periodic=(0, 0, 1))


def _fluid_area(spec, n, hi=(1.0, 1.0, None)):
    g = _geom(n, hi)
    gm = ocfd.eb_geometry(spec, g, (0, 0, 0), (n[0] - 1, n[1] - 1, 0), 0)
    return gm["vfrac"][0].sum() * g.dx[0] * g.dx[1], gm, g


def test_plane_and_box_areas_are_exact():
    """Straight boundaries: the cut polygons tile the fluid region exactly."""
    # ramp of Exec/4_WEDGE/inputs (plane point (2, 0.2), normal (0.5, -1)) on [0,3]x[0,1]
    sp = ocfd.eb_plane((2.0, 0.2), (0.5, -1.0))
    A, _, _ = _fluid_area(sp, (96, 32, 1), (3.0, 1.0))
    # body: y < 0.2 + 0.5 (x - 2) for x in [1.6, 3]: a triangle up to x=3 where y = 0.7
    body = 0.5 * (3.0 - 1.6) * 0.7
    assert abs(A - (3.0 - body)) < 1e-12
    # step of Exec/3_FSTEP/inputs shifted off the grid lines: body [0.61, 4] x [-1, 0.213]
    sb = ocfd.eb_box((0.61, -1.0), (4.0, 0.213))
    A, gm, _ = _fluid_area(sb, (120, 40, 1), (3.0, 1.0))
    # the one cell holding the body's corner (x 0.6..0.625, y 0.2..0.225) is cut
    # by the chord between its edge roots, which adds the triangle
    # 0.015 x 0.013 / 2 to the fluid
    assert abs(A - (3.0 - (3.0 - 0.61) * 0.213 + 0.5 * 0.015 * 0.013)) < 1e-12


def test_cylinder_area_converges_second_order():
    errs = []
    for n in (64, 128, 256):
        sp = ocfd.eb_cylinder(0.5, 0.5, 0.2)
        A, _, _ = _fluid_area(sp, (n, n, 1))
        errs.append(abs(A - (1.0 - math.pi * 0.04)))
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs


def test_cut_polygons_close():
    """Each cut cell's faces and its straight boundary segment form a closed
    polygon: the wall vector ((apx_lo - apx_hi) dy, (apy_lo - apy_hi) dx) is
    normal to the chord, hence radial at the chord's midpoint (bcent)."""
    n = 64
    sp = ocfd.eb_cylinder(0.5, 0.5, 0.2)
    g = _geom((n, n, 1))
    gm = ocfd.eb_geometry(sp, g, (0, 0, 0), (n - 1, n - 1, 0), 1)   # one ghost ring: hi faces
    f = gm["flag"][1]
    dx = g.dx[0]
    cut = 0
    for j in range(1, n + 1):
        for i in range(1, n + 1):
            if (f[j, i] & 3) != 1:
                continue
            cut += 1
            wx = (gm["apx"][1, j, i] - gm["apx"][1, j, i + 1]) * dx
            wy = (gm["apy"][1, j, i] - gm["apy"][1, j + 1, i]) * dx
            bx = (i - 1 + 0.5 + gm["bcent"][0, 1, j, i]) * dx - 0.5
            by = (j - 1 + 0.5 + gm["bcent"][1, 1, j, i]) * dx - 0.5
            cross = wx * by - wy * bx
            assert abs(cross) <= 1e-9 * math.hypot(wx, wy) * math.hypot(bx, by), (i, j)
            # the wall vector points into the body (towards the centre)
            assert wx * bx + wy * by < 0
            assert 0.0 < gm["vfrac"][1, j, i] < 1.0 or gm["vfrac"][1, j, i] == 1.0
    assert cut > 40


def test_flags_connectivity_and_types():
    n = 32
    sp = ocfd.eb_cylinder(0.5, 0.5, 0.21)
    g = _geom((n, n, 1))
    gm = ocfd.eb_geometry(sp, g, (0, 0, 0), (n - 1, n - 1, 0), 1)
    f, vf, ax, ay = gm["flag"][1], gm["vfrac"][1], gm["apx"][1], gm["apy"][1]

    def conn(w, i, j, k):
        return (int(w) >> (5 + (i + 1) + 3 * (j + 1) + 9 * (k + 1))) & 1

    for j in range(1, n + 1):
        for i in range(1, n + 1):
            w = int(f[j, i])
            t = w & 3
            if vf[j, i] == 0.0:
                assert t == 3 and (w >> 5) == 0
                continue
            assert (t == 0) == (vf[j, i] == 1.0 and ax[j, i] == 1 and ax[j, i + 1] == 1 and ay[j, i] == 1 and ay[j + 1, i] == 1)
            assert conn(w, 0, 0, 0) and conn(w, 0, 0, 1) and conn(w, 0, 0, -1)
            assert conn(w, -1, 0, 0) == (ax[j, i] > 0 and vf[j, i - 1] > 0)
            assert conn(w, 1, 0, 0) == (ax[j, i + 1] > 0 and vf[j, i + 1] > 0)
            assert conn(w, 0, -1, 0) == (ay[j, i] > 0 and vf[j - 1, i] > 0)
            assert conn(w, 0, 1, 0) == (ay[j + 1, i] > 0 and vf[j + 1, i] > 0)
            if t == 0 and (f[j - 1:j + 2, i - 1:i + 2] & 3 == 0).all():
                assert (w >> 5) == (1 << 27) - 1      # all 27 neighbours


def test_small_cells_become_covered():
    n = 32
    sp = ocfd.eb_cylinder(0.5, 0.5, 0.2, small_volfrac=0.2)
    gm = ocfd.eb_geometry(sp, _geom((n, n, 1)), (0, 0, 0), (n - 1, n - 1, 0), 0)
    vf = gm["vfrac"][0]
    assert not ((vf > 0) & (vf < 0.2)).any()
    assert ((gm["flag"][0] & 3) == 3).sum() == (vf == 0).sum()


def test_channel_has_all_box_kinds(orc):
    """The channel's box list holds regular, cut and covered boxes; the oracle
    classifies them as compute_dSdt does (Source/Advance.cpp:213,231)."""
    boxes, g, kind, sp = cases.eb_channel(256, 64, 2, max_grid=8)
    kinds = []
    for lo, hi in boxes:
        glo, ghi = orc.grown(lo, hi)
        gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
        e = orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"])
        kinds.append(orc.eb_box_kind(lo, hi, e))
    assert set(kinds) == {0, 1, 2}


def _prim(rho, u, v, w, p, p_=None):
    par = p_
    T = p / (rho * par.Rg)
    c = math.sqrt(par.eos_gamma * par.Rg * T)
    return [rho, u, v, w, p, T, c]


def test_van_leer_split_is_consistent(orc):
    """f+(q) + f-(q) equals the Euler flux for every Mach number, and the
    split is pure upwind for |M| >= 1."""
    par = orc.parm()
    rng = np.random.default_rng(0)
    for d in range(3):
        q = []
        for _ in range(200):
            rho = rng.uniform(0.1, 3)
            p = rng.uniform(0.1, 3) * 1e5
            c = math.sqrt(par.eos_gamma * p / rho)
            vel = rng.uniform(-2.5, 2.5, 3) * c
            q.append(_prim(rho, *vel, p, par))
        q = np.array(q)
        F = orc.eb_flux(d, q, q, par)
        g = par.eos_gamma
        un = q[:, 1 + d]
        E = q[:, 4] / (g - 1) + 0.5 * q[:, 0] * (q[:, 1] ** 2 + q[:, 2] ** 2 + q[:, 3] ** 2)
        exact = np.stack([q[:, 0] * un,
                          q[:, 0] * un * q[:, 1] + (d == 0) * q[:, 4],
                          q[:, 0] * un * q[:, 2] + (d == 1) * q[:, 4],
                          q[:, 0] * un * q[:, 3] + (d == 2) * q[:, 4],
                          un * (E + q[:, 4])], axis=1)
        scale = np.abs(exact).max(axis=1, keepdims=True)
        assert np.abs(F - exact).max() <= 1e-12 * scale.max()
        # both states supersonic in +d: the flux is the left state's alone
        sup = np.nonzero(q[:, 1 + d] / q[:, 6] >= 1.0)[0]
        assert len(sup) > 5
        F2 = orc.eb_flux(d, q[sup], q[sup[::-1]], par)
        assert np.allclose(F2, exact[sup], rtol=1e-13, atol=0)


def _eb_level_fabs(orc, boxes, g, sp):
    out = []
    for lo, hi in boxes:
        glo, ghi = orc.grown(lo, hi)
        gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
        out.append(orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"]))
    return out


def _og(orc, g):
    import ctypes
    og = orc.OGeom()
    ctypes.memmove(ctypes.byref(og), ctypes.byref(g), ctypes.sizeof(og))
    return og


@pytest.mark.parametrize("visc", [0, 1])
def test_gas_at_rest_stays_at_rest(orc, visc):
    """Uniform pressure, zero velocity: every face flux is p n and the wall
    term (apx_lo - apx_hi) p closes each cut cell, so dSdt = 0 everywhere
    (pins the area-fraction divergence and the wall pressure term)."""
    boxes, g, kind, sp = cases.eb_channel(128, 32, 2, max_grid=16)
    par = orc.parm()
    ebs = _eb_level_fabs(orc, boxes, g, sp)
    og = _og(orc, g)
    for (lo, hi), e in zip(boxes, ebs):
        if orc.eb_box_kind(lo, hi, e) != 1:
            continue
        S = orc.Fab(*orc.grown(lo, hi))
        S.a[0] = 1.2
        S.a[4] = 1e5 / (par.eos_gamma - 1.0)
        D = orc.compute_dsdt_box_eb(lo, hi, S, par, [1 / g.dx[0]] * 3, e, do_visc=visc)
        scale = 1e5 / g.dx[0]
        assert np.abs(D).max() <= 1e-10 * scale, np.abs(D).max() / scale


def test_eb_path_conserves_on_periodic_box(orc):
    """One periodic box around a small cylinder goes through the EB path: the
    face fluxes telescope, and the inviscid wall flux carries momentum only,
    so total mass and energy change by rounding only."""
    n = 32
    g = ocfd.make_geom((n, n, 4), (0, 0, 0), (1, 1, 4.0 / n), periodic=(1, 1, 1))
    sp = ocfd.eb_cylinder(0.5, 0.5, 0.1)
    boxes = [((0, 0, 0), (n - 1, n - 1, 3))]
    par = orc.parm()
    og = _og(orc, g)
    ebs = _eb_level_fabs(orc, boxes, g, sp)
    assert orc.eb_box_kind(*boxes[0], ebs[0]) == 1
    rng = np.random.default_rng(3)
    A = orc.new_level_fabs(boxes)
    lo, hi = boxes[0]
    v = A[0].valid(lo, hi)
    v[0] = rng.uniform(1.0, 1.5, v[0].shape)
    v[1] = rng.uniform(-0.2, 0.2, v[0].shape) * v[0]
    v[2] = rng.uniform(-0.2, 0.2, v[0].shape) * v[0]
    v[4] = rng.uniform(2.0, 3.0, v[0].shape) + 0.5 * (v[1] ** 2 + v[2] ** 2) / v[0]
    assert orc.fill_ghost(boxes, A, og, 0.0) == 0
    D = orc.compute_dsdt_box_eb(lo, hi, A[0], par, [1 / g.dx[0]] * 3, ebs[0])
    vf = ebs[0].arrs[0][4:-4, 4:-4, 4:-4]
    tot = (D * vf[None]).reshape(5, -1).sum(axis=1)
    mag = (np.abs(D) * vf[None]).reshape(5, -1).sum(axis=1)
    assert abs(tot[0]) < 1e-13 * mag[0] * 1e3 and abs(tot[4]) < 1e-13 * mag[4] * 1e3, (tot, mag)
    assert abs(tot[1]) > 1e-6 * mag[1]          # the wall pressure (not uniform) pushes


def test_free_stream_away_from_the_body(orc):
    """Uniform flow: in an EB box, cells whose stencils see only regular cells
    keep dSdt = 0 exactly (f+ + f- of identical states telescopes)."""
    boxes, g, kind, sp = cases.eb_channel(128, 32, 2, max_grid=16)
    par = orc.parm()
    ebs = _eb_level_fabs(orc, boxes, g, sp)
    st = orc.inflow_state()
    checked = 0
    for (lo, hi), e in zip(boxes, ebs):
        if orc.eb_box_kind(lo, hi, e) != 1:
            continue
        S = orc.Fab(*orc.grown(lo, hi))
        for n in range(5):
            S.a[n] = st[n]
        D = orc.compute_dsdt_box_eb(lo, hi, S, par, [1 / g.dx[0]] * 3, e)
        f = e.flag
        reg = (f & 3) == 0
        # regular within 3 cells in x and y
        ok = np.ones_like(reg)
        for dj in range(-3, 4):
            for di in range(-3, 4):
                ok &= np.roll(np.roll(reg, dj, axis=1), di, axis=2)
        ok = ok[4:-4, 4:-4, 4:-4]
        assert ok.any()
        assert (D[:, ok] == 0.0).all()
        assert np.abs(D[:, ~ok]).max() > 0
        checked += 1
    assert checked > 0
