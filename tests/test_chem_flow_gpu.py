"""GPU parity of the multispecies (CHEM) flow path against the oracle:
initial data and species ghost fill (bit-exact), dSdt of state and species
(inviscid and viscous, 1e-10 relative), and Strang-split steps of the
Exec/5_CHEMSOD shock tube (RK2 flow between chemistry half steps, species
renormalisation), 1e-10 relative."""
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
RU = 8.31446261815324


def _level(case):
    import ocfd
    boxes, g, kind = case
    return ocfd.Level(boxes, g), boxes, g, kind


def test_init_and_species_fill_match_oracle(torch_cuda, orc):
    import ocfd
    from ocfd import cases
    lev, boxes, g, kind = _level(cases.chemsod(96, 4, 4, max_grid=32))
    p = ocfd.default_parm()
    S, P = lev.new_mf(), lev.new_spec_mf()
    lev.init_species_state(kind, S, P, p)
    og = gu.to_ogeom(g)
    fa, fs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    orc.init_chemsod(boxes, fa, fs, og)
    got_s = gu.download(lev, S)
    got_p = [lev.fab_view(P, b, 4, 9).cpu().numpy() for b in range(lev.nbox)]
    for b in range(lev.nbox):
        assert np.array_equal(got_s[b], fa[b].a), b
        assert np.array_equal(got_p[b], fs[b].a), b
    # species ghost fill: random partial densities
    rng = np.random.default_rng(4)
    for f in fs:
        f.a[...] = rng.uniform(0, 1, f.a.shape)
    for b, f in enumerate(fs):
        import torch
        lev.fab_view(P, b, 4, 9).copy_(torch.from_numpy(f.a))
    lev.fill_boundary_spec(P, 0.0)
    assert orc.fill_ghost_spec(boxes, fs, og, 0.0) == 0
    for b, f in enumerate(fs):
        assert np.array_equal(lev.fab_view(P, b, 4, 9).cpu().numpy(), f.a), b


def _mixture_state(orc, n, shape=None):
    """A smooth reacting-mixture state on a periodic box of n^3 (or shape
    (nz, ny, nx)) valid cells."""
    shape = shape or (n, n, n)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*(np.arange(m) for m in shape), indexing="ij")
    Y = np.zeros((9,) + tuple(shape))
    Y[0] = 0.02 + 0.01 * np.sin(2 * np.pi * x / nx)
    Y[1] = 0.22
    Y[5] = 1e-3 * (1 + np.cos(2 * np.pi * z / nz))
    Y[8] = 1 - Y.sum(0)
    T = 900 + 100 * np.cos(2 * np.pi * y / ny)
    pr = 101325 * (1 + 0.1 * np.sin(2 * np.pi * z / nz))
    u = 10 * np.sin(2 * np.pi * y / ny)
    v = 5 * np.cos(2 * np.pi * x / nx)
    w = 3 * np.sin(2 * np.pi * x / nx)
    mw = orc.chem_tables()[0]
    rho = pr / (RU * T * (Y / mw[:, None, None, None]).sum(0))
    rhoi = rho * Y
    e = np.array([orc.thermo("CKUBMS", T.flat[c], rhoi.reshape(9, -1)[:, c]) for c in range(T.size)]).reshape(T.shape)
    U = np.stack([rho, rho * u, rho * v, rho * w, e + 0.5 * rho * (u * u + v * v + w * w)])
    return U, rhoi


@pytest.mark.parametrize("visc", [0, 1])
def test_compute_dsdt_multi(torch_cuda, orc, visc):
    import torch
    import ocfd
    from ocfd import cases
    n = 16
    boxes = cases.chop((n, n, n), (8, 16, 16))
    g = ocfd.make_geom((n, n, n), (0, 0, 0), (0.01, 0.01, 0.01), periodic=(1, 1, 1))
    lev = ocfd.Level(boxes, g)
    og = gu.to_ogeom(g)
    U, rhoi = _mixture_state(orc, n)
    fa, fs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    for (lo, hi), f, fsp in zip(boxes, fa, fs):
        sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
        f.valid(lo, hi)[:] = U[sl]
        fsp.valid(lo, hi)[:] = rhoi[sl]
    assert orc.fill_ghost(boxes, fa, og, 0.0) == 0
    assert orc.fill_ghost_spec(boxes, fs, og, 0.0) == 0
    S, P = lev.new_mf(), lev.new_spec_mf()
    gu.upload(lev, S, fa)
    for b, f in enumerate(fs):
        lev.fab_view(P, b, 4, 9).copy_(torch.from_numpy(f.a))
    D, Dp = lev.new_mf(ngrow=0), lev.new_mf(ngrow=0, ncomp=9)
    p = ocfd.default_parm()
    lev.compute_dsdt_multi(S, P, D, Dp, p, ocfd.VISC if visc else 0)
    got = gu.download(lev, D, ngrow=0)
    got_p = [lev.fab_view(Dp, b, 0, 9).cpu().numpy() for b in range(lev.nbox)]
    ref, ref_p = [], []
    for (lo, hi), f, fsp in zip(boxes, fa, fs):
        d, dp = orc.compute_dsdt_multi_box(lo, hi, f, fsp, orc.parm(), [1 / g.dx[0]] * 3, visc)
        ref.append(d)
        ref_p.append(dp)
    e1 = gu.level_max_rel(got, ref)
    e2 = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(got_p, ref_p))
    print("multi dSdt visc", visc, "state", e1, "species", e2)
    assert e1 < TOL64 and e2 < TOL64


@pytest.mark.parametrize("visc", [0, 1])
def test_chemsod_steps(torch_cuda, orc, visc):
    """Strang-split steps (level 0: 4 chemistry substeps on each side of the
    RK2 flow step) of the H2/O2 shock tube with the hot driver."""
    import ocfd
    from ocfd import cases
    lev, boxes, g, kind = _level(cases.chemsod(128, 4, 4, max_grid=64))
    p = ocfd.default_parm()
    A, As, B, Bs = lev.new_mf(), lev.new_spec_mf(), lev.new_mf(), lev.new_spec_mf()
    lev.init_species_state(kind, A, As, p)
    flags = ocfd.VISC if visc else 0
    steps, t, dt = ocfd.run_chem(lev, A, As, B, Bs, p, cfl=1.0, nsteps=6, flags=flags)
    og = gu.to_ogeom(g)
    fa, fs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    fb, fbs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    orc.init_chemsod(boxes, fa, fs, og)
    tt = 0.0
    for _ in range(6):
        assert orc.advance_chem(boxes, fa, fs, fb, fbs, og, orc.parm(), tt, 1e-8, 0, 1, visc, 1) == 0
        tt += 1e-8
    got = gu.download(lev, A)
    got_p = [lev.fab_view(As, b, 4, 9).cpu().numpy() for b in range(lev.nbox)]
    e1 = gu.level_max_rel([x[:, 4:-4, 4:-4, 4:-4] for x in got], [f.a[:, 4:-4, 4:-4, 4:-4] for f in fa])
    e2 = max(np.abs(a[:, 4:-4, 4:-4, 4:-4] - f.a[:, 4:-4, 4:-4, 4:-4]).max() / np.abs(f.a).max()
             for a, f in zip(got_p, fs))
    print("chemsod visc", visc, "state", e1, "species", e2)
    assert steps == 6 and abs(t - 6e-8) < 1e-20
    assert e1 < TOL64 and e2 < TOL64
    # the driver ignites: water appears in the hot cells
    assert got_p[0][2, 4:-4, 4:-4, 4:8].max() > 0


def _eb_fabs(orc, boxes, g, sp):
    import ocfd
    out = []
    for lo, hi in boxes:
        glo, ghi = orc.grown(lo, hi)
        gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
        out.append(orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"]))
    return out


@pytest.mark.parametrize("visc", [0, 1])
def test_ebchem_level(torch_cuda, orc, visc):
    """A reacting mixture around a cylinder: regular boxes through the
    multispecies WENO path, cut boxes through the EB-CHEM path (species only),
    covered boxes frozen; dSdt and two RK2 steps against the oracle."""
    import torch
    import ocfd
    from ocfd import cases
    nx, ny, nz = 64, 64, 4
    boxes = cases.chop((nx, ny, nz), (8, 8, 4))
    g = ocfd.make_geom((nx, ny, nz), (0, 0, 0), (0.01, 0.01, 0.01 * nz / nx), periodic=(1, 1, 1))
    sp = ocfd.eb_cylinder(0.005, 0.005, 0.0021, small_volfrac=0.05)
    lev = ocfd.Level(boxes, g, eb=sp)
    kinds = lev.box_types()
    assert set(kinds) == {0, 1, 2}
    og = gu.to_ogeom(g)
    U, rhoi = _mixture_state(orc, 0, (nz, ny, nx))
    fa, fs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    for (lo, hi), f, fsp in zip(boxes, fa, fs):
        sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
        f.valid(lo, hi)[:] = U[sl]
        fsp.valid(lo, hi)[:] = rhoi[sl]
    assert orc.fill_ghost(boxes, fa, og, 0.0) == 0
    assert orc.fill_ghost_spec(boxes, fs, og, 0.0) == 0
    S, P = lev.new_mf(), lev.new_spec_mf()
    gu.upload(lev, S, fa)
    for b, f in enumerate(fs):
        lev.fab_view(P, b, 4, 9).copy_(torch.from_numpy(f.a))
    D, Dp = lev.new_mf(ngrow=0), lev.new_mf(ngrow=0, ncomp=9)
    p = ocfd.default_parm()
    flags = ocfd.VISC if visc else 0
    lev.compute_dsdt_multi(S, P, D, Dp, p, flags)
    got = gu.download(lev, D, ngrow=0)
    got_p = [lev.fab_view(Dp, b, 0, 9).cpu().numpy() for b in range(lev.nbox)]
    ebs = _eb_fabs(orc, boxes, g, sp)
    for kd in (0, 1, 2):
        sel = [b for b in range(len(boxes)) if kinds[b] == kd]
        ref, ref_p = [], []
        for b in sel:
            lo, hi = boxes[b]
            d, dp = orc.compute_dsdt_multi_box_eb(lo, hi, fa[b], fs[b], orc.parm(), [1 / g.dx[0]] * 3, ebs[b], visc)
            ref.append(d)
            ref_p.append(dp)
        if kd != 0:
            assert all(np.abs(got[b]).max() == 0.0 for b in sel)     # flow dSdt 0 in cut and covered boxes
        if kd == 2:
            assert all(np.abs(got_p[b]).max() == 0.0 for b in sel)
            continue
        if kd == 0:
            assert gu.level_max_rel([got[b] for b in sel], ref) < TOL64
        e2 = max(np.abs(got_p[b] - r).max() for b, r in zip(sel, ref_p)) / max(np.abs(r).max() for r in ref_p)
        print("ebchem kind", kd, "species", e2)
        assert e2 < TOL64
    # two RK2 steps of the level
    B, Bp = lev.new_mf(), lev.new_spec_mf()
    fb, fbs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    orc.set_eb(ebs)
    try:
        for st in range(2):
            lev.advance_multi(S, P, B, Bp, st * 1e-8, 1e-8, 1, p, flags)
            assert orc.flow_advance_multi(boxes, fa, fs, fb, fbs, og, orc.parm(), st * 1e-8, 1e-8, 1, visc) == 0
    finally:
        orc.set_eb(None)
    got = gu.download(lev, S)
    got_p = [lev.fab_view(P, b, 4, 9).cpu().numpy() for b in range(lev.nbox)]
    e1 = gu.level_max_rel([x[:, 4:-4, 4:-4, 4:-4] for x in got], [f.a[:, 4:-4, 4:-4, 4:-4] for f in fa])
    e2 = max(np.abs(a[:, 4:-4, 4:-4, 4:-4] - f.a[:, 4:-4, 4:-4, 4:-4]).max() / np.abs(f.a).max()
             for a, f in zip(got_p, fs))
    print("ebchem steps state", e1, "species", e2)
    assert e1 < TOL64 and e2 < TOL64
