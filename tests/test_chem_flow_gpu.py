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


def _mixture_state(orc, n):
    """A smooth reacting-mixture state on a periodic n^3 box (valid cells)."""
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    Y = np.zeros((9, n, n, n))
    Y[0] = 0.02 + 0.01 * np.sin(2 * np.pi * x / n)
    Y[1] = 0.22
    Y[5] = 1e-3 * (1 + np.cos(2 * np.pi * z / n))
    Y[8] = 1 - Y.sum(0)
    T = 900 + 100 * np.cos(2 * np.pi * y / n)
    pr = 101325 * (1 + 0.1 * np.sin(2 * np.pi * z / n))
    u = 10 * np.sin(2 * np.pi * y / n)
    v = 5 * np.cos(2 * np.pi * x / n)
    w = 3 * np.sin(2 * np.pi * x / n)
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
