"""Pins of the oracle's embedded-boundary point kernels by facts their
mathematics fixes, on a cut box around a cylinder (the library's own cut-cell
geometry, whose flags and connectivity test_eb_cpu.py pins):

- eb_recon (Source/EB/EBkernels.H:10-71): half minmod slopes reproduce linear
  data exactly where the connectivity allows them and fall back to the cell
  value exactly where it does not; the limited slope of a kink is the smaller
  one-sided slope; every face value lies between its two cells (TVD).
- eb_compute_visc (Source/EB/EBdiffusion.H:8-203, weights {0, 1, 0.5} of
  Source/EB/AdvanceBox.cpp:40-43): a linear shear gives the exact stress,
  one-sided where a tangential neighbour is not connected; a cell with no
  connected tangential neighbour contributes nothing to the tangential
  derivative.
- eb_wallflux (Source/EB/EBkernels.H:173-214): no mass or energy crosses the
  wall, the momentum flux is p times the area-fraction normal at rest, rotates
  with the flow and the wall, and scales with the area vector.

The +-2 connectivity query of eb_recon reads the flag bit the linear AMReX
EBCellFlag layout gives that offset (AMReX is not vendored, so that choice is
not pinned here; the flag layout for +-1 is)."""
import math

import numpy as np
import pytest

import ocfd

N = 32


def _setup(orc, nz=4, radius=0.21):
    g = ocfd.make_geom((N, N, nz), (0, 0, 0), (1, 1, nz / N), periodic=(0, 0, 1))
    sp = ocfd.eb_cylinder(0.5, 0.5, radius)
    lo, hi = (0, 0, 0), (N - 1, N - 1, nz - 1)
    glo, ghi = orc.grown(lo, hi)
    gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
    eb = orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"])
    return g, lo, hi, glo, ghi, eb


def _bit(w, o):
    """EBCellFlag bit of neighbour offset o (AMReX linear layout)."""
    b = 5 + (o[0] + 1) + 3 * (o[1] + 1) + 9 * (o[2] + 1)
    return (int(w) >> b) & 1


def _covered(w):
    return (int(w) & 3) == 3


E = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_eb_recon_linear_exact_where_connected(orc, d):
    g, lo, hi, glo, ghi, eb = _setup(orc)
    q = orc.Fab(glo, ghi, ncomp=2)
    nz, ny, nx = q.a.shape[1:]
    idx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")[2 - d]
    q.a[0] = 1.0 + 0.25 * idx                       # linear along d, exact in binary
    q.a[1] = np.random.default_rng(d).uniform(-1, 1, q.a[1].shape)
    ql, qr = orc.eb_recon_box(d, lo, hi, q, eb)
    e = E[d]
    counts = {"both": 0, "first_order": 0}
    fshape = ql.shape[1:]
    for k in range(fshape[0]):
        for j in range(fshape[1]):
            for i in range(fshape[2]):
                gi, gj, gk = i + 4, j + 4, k + 4                 # face (i,j,k) in grown indices: high cell
                lowc = (gk - e[2], gj - e[1], gi - e[0])
                w = eb.flag[gk, gj, gi]
                qlo, qhi = q.a[0][lowc], q.a[0][gk, gj, gi]
                face = 0.5 * (qlo + qhi)                       # exact face value of linear data
                c1 = _bit(w, tuple(-x for x in e))
                c2 = _bit(w, tuple(-2 * x for x in e))
                cp = _bit(w, e)
                exp_l = face if (c1 and c2) else qlo
                exp_r = face if (c1 and cp) else qhi
                assert ql[0, k, j, i] == exp_l and qr[0, k, j, i] == exp_r, (d, k, j, i)
                counts["both" if (c1 and c2 and cp) else "first_order"] += 1
                # TVD: every face value within its two cells' range
                a, b = q.a[1][lowc], q.a[1][gk, gj, gi]
                assert min(a, b) <= ql[1, k, j, i] <= max(a, b)
                assert min(a, b) <= qr[1, k, j, i] <= max(a, b)
    assert counts["both"] > 0 and counts["first_order"] > 0, counts


def test_eb_recon_limits_a_kink_to_the_smaller_slope(orc):
    """Piecewise-linear data along x, slope 1/2 then 2 past x0 (exact in
    binary): at the face after the kink the left state takes the smaller
    one-sided slope and the right state the (equal) slopes of its side; at
    an extremum both states fall back to the cell values."""
    g, lo, hi, glo, ghi, eb = _setup(orc)
    q = orc.Fab(glo, ghi, ncomp=2)
    nz, ny, nx = q.a.shape[1:]
    x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")[2].astype(float)
    x0 = 8.0
    q.a[0] = np.where(x <= x0, 0.5 * x, 0.5 * x0 + 2.0 * (x - x0))
    q.a[1] = np.abs(x - x0)                          # a minimum at x0
    ql, qr = orc.eb_recon_box(0, lo, hi, q, eb)
    j = 4                                            # a y row far from the cylinder: regular cells
    assert ((eb.flag[4:-4, j + 4, :] & 3) == 0).all()
    fi = int(x0) + 1 - 4                             # face between grown cells x0 and x0+1
    for k in range(ql.shape[1]):
        assert ql[0, k, j, fi] == q.a[0][k + 4, j + 4, int(x0)] + 0.5 * 0.5   # min(|1/2|, |2|)
        assert qr[0, k, j, fi] == q.a[0][k + 4, j + 4, int(x0) + 1] - 0.5 * 2.0
        fe = int(x0) - 4                             # face whose high cell is the minimum
        # low cell x0-1, high cell x0 (the minimum): right state's slopes +1 / -1 -> 0
        assert qr[1, k, j, fe] == q.a[1][k + 4, j + 4, int(x0)]
        # face after the minimum: left state at the minimum has slopes +1 / -1 -> 0
        assert ql[1, k, j, fe + 1] == q.a[1][k + 4, j + 4, int(x0)]


def _mu(parm, T):
    return parm.C_s * T ** 1.5 / (T + parm.T_s)       # Sutherland's law


@pytest.mark.parametrize("d", [0, 1])
def test_eb_visc_linear_shear_one_sided(orc, d):
    """u = alpha y, v = beta x (uniform rho, p, T): on a d-face the shear
    stress is mu (du_d/dt + du_t/dd) with the tangential derivative exact
    from each cell that has a connected neighbour along it (one-sided where
    only one is), so tau = mu (beta + alpha n/2) (x faces) or mu (alpha +
    beta n/2) (y faces) with n the number of the face's two cells having
    one; normal stress, tau_xz and the heat flux vanish."""
    g, lo, hi, glo, ghi, eb = _setup(orc)
    par = orc.parm()
    dx = g.dx[0]
    alpha, beta, T = 3.0e3, -1.25e3, 300.0
    q = orc.Fab(glo, ghi, ncomp=7)
    nz, ny, nx = q.a.shape[1:]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    xc, yc = (x - 4 + 0.5) * dx, (y - 4 + 0.5) * dx
    q.a[0], q.a[1], q.a[2], q.a[3] = 1.2, alpha * yc, beta * xc, 0.0
    q.a[4], q.a[5], q.a[6] = 1e5, T, 347.0
    t = 1 - d                                     # the in-plane tangential direction
    e, et = E[d], E[t]
    # cells with no connected neighbour along t (weight 0): clear both bits in
    # a few regular cells away from the body (eb_visc reads only these bits)
    for (gk, gj, gi) in ((5, 8, 8), (6, 30, 12), (5, 10, 29)):
        for o in (et, tuple(-x for x in et)):
            eb.flag[gk, gj, gi] &= ~np.uint32(1 << (5 + (o[0] + 1) + 3 * (o[1] + 1) + 9 * (o[2] + 1)))
    F = orc.eb_visc_box(d, lo, hi, q, eb, [1 / dx] * 3, par)
    mu = _mu(par, T)
    seen = {0: 0, 1: 0, 2: 0}
    onesided = 0
    for k in range(F.shape[1]):
        for j in range(F.shape[2]):
            for i in range(F.shape[3]):
                gk, gj, gi = k + 4, j + 4, i + 4
                wh, wl = eb.flag[gk, gj, gi], eb.flag[gk - e[2], gj - e[1], gi - e[0]]
                if _covered(wh) or _covered(wl):
                    continue
                has = [(_bit(w, et) or _bit(w, tuple(-x for x in et))) for w in (wh, wl)]
                onesided += sum(_bit(w, et) != _bit(w, tuple(-x for x in et)) for w in (wh, wl))
                n = int(has[0]) + int(has[1])
                seen[n] += 1
                tang = (alpha if d == 0 else beta) * n / 2
                normal = beta if d == 0 else alpha
                tau = mu * (tang + normal)
                got = F[:, k, j, i]
                assert abs(got[2 if d == 0 else 1] + tau) <= 1e-12 * abs(mu * (abs(alpha) + abs(beta))), (k, j, i)
                assert got[0] == 0.0
                assert abs(got[1 + d]) <= 1e-12 * mu * abs(alpha)      # normal stress: div u = 0
                assert abs(got[3]) <= 1e-12 * mu * abs(alpha)
                vs = q.a[1 + t][gk, gj, gi] + q.a[1 + t][gk - e[2], gj - e[1], gi - e[0]]
                assert abs(got[4] + 0.5 * vs * tau) <= 1e-12 * abs(0.5 * vs * tau) + 1e-9
    assert seen[2] > 0 and seen[1] > 0 and onesided > 0, (seen, onesided)


def test_eb_wallflux_conserves_rotates_and_scales(orc):
    par = orc.parm()
    rng = np.random.default_rng(11)
    n = 300
    d = rng.uniform(0.1, 3.0, n)
    p = rng.uniform(1e4, 1e6, n)
    c = np.sqrt(par.eos_gamma * p / d)
    vel = rng.uniform(-300, 300, (n, 3))
    qw = np.column_stack([d, vel, p, p / (d * par.Rg), c])
    area = rng.uniform(0, 1, (n, 6))
    F = orc.eb_wallflux(qw, area, par)
    # (a) no mass or energy through the wall, for any state
    assert (F[:, 0] == 0).all() and (F[:, 4] == 0).all()
    # (b) at rest: momentum flux = p (a_lo - a_hi) per direction
    q0 = qw.copy()
    q0[:, 1:4] = 0
    F0 = orc.eb_wallflux(q0, area, par)
    nvec = np.column_stack([area[:, 0] - area[:, 1], area[:, 2] - area[:, 3], area[:, 4] - area[:, 5]])
    assert np.array_equal(F0[:, 1:4], p[:, None] * nvec)
    # (c) rotating flow and wall about z rotates the momentum flux
    th = rng.uniform(0, 2 * math.pi, n)
    cs, sn = np.cos(th), np.sin(th)
    rot = lambda a, b: (cs * a - sn * b, sn * a + cs * b)
    qr_ = qw.copy()
    qr_[:, 1], qr_[:, 2] = rot(qw[:, 1], qw[:, 2])
    nx, ny = rot(nvec[:, 0], nvec[:, 1])
    ar = area.copy()
    ar[:, 0], ar[:, 2] = ar[:, 1] + nx, ar[:, 3] + ny
    Fr = orc.eb_wallflux(qr_, ar, par)
    fx, fy = rot(F[:, 1], F[:, 2])
    scale = np.abs(F[:, 1:4]).max(axis=1)
    assert (np.abs(Fr[:, 1] - fx) <= 1e-12 * scale).all() and (np.abs(Fr[:, 2] - fy) <= 1e-12 * scale).all()
    assert (np.abs(Fr[:, 3] - F[:, 3]) <= 1e-12 * scale).all()
    # (d) the flux is homogeneous of degree one in the area vector
    s = 2.5
    asc = area.copy()
    asc[:, 0], asc[:, 2], asc[:, 4] = area[:, 1] + s * nvec[:, 0], area[:, 3] + s * nvec[:, 1], area[:, 5] + s * nvec[:, 2]
    Fs = orc.eb_wallflux(qw, asc, par)
    assert (np.abs(Fs[:, 1:4] - s * F[:, 1:4]) <= 1e-12 * s * scale[:, None]).all()
