"""GPU parity of the embedded-boundary (cut-cell) box path against the oracle:
box classification, dSdt of every box kind (regular WENO boxes, cut boxes
through the MUSCL / Van Leer path, covered boxes), and whole SSP-RK3 steps of
the supersonic channel past a cylinder.  FP64 tolerance 1e-10 relative."""
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _eb_fabs(orc, boxes, g, sp):
    import ocfd
    out = []
    for lo, hi in boxes:
        glo, ghi = orc.grown(lo, hi)
        gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
        out.append(orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"]))
    return out


def _valid_flags(boxes, g, sp):
    import ocfd
    return [ocfd.eb_geometry(sp, g, lo, hi, 0)["flag"] for lo, hi in boxes]


def test_box_kinds_match_oracle(torch_cuda, orc):
    import ocfd
    from ocfd import cases
    boxes, g, kind, sp = cases.eb_channel(256, 64, 2, max_grid=8)
    lev = ocfd.Level(boxes, g, eb=sp)
    ebs = _eb_fabs(orc, boxes, g, sp)
    want = [orc.eb_box_kind(lo, hi, e) for (lo, hi), e in zip(boxes, ebs)]
    assert lev.box_types() == want
    assert set(want) == {0, 1, 2}


def test_init_matches_oracle(torch_cuda, orc):
    import ocfd
    from ocfd import cases
    boxes, g, kind, sp = cases.eb_channel(128, 32, 2, max_grid=16)
    lev = ocfd.Level(boxes, g, eb=sp)
    p = ocfd.default_parm()
    A = lev.new_mf()
    lev.init_state(kind, A, p)
    fa = orc.new_level_fabs(boxes)
    orc.init_eb_inflow(boxes, fa, _valid_flags(boxes, g, sp), orc.parm())
    got = gu.download(lev, A)
    for x, f in zip(got, fa):
        assert np.array_equal(x[:, 4:-4, 4:-4, 4:-4], f.a[:, 4:-4, 4:-4, 4:-4])


@pytest.mark.parametrize("visc,grav,redist", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_compute_dsdt_eb_level(torch_cuda, orc, visc, grav, redist):
    """dSdt of a perturbed channel state, every box kind, against the oracle."""
    import ocfd
    from ocfd import cases
    boxes, g, kind, sp = cases.eb_channel(256, 64, 2, max_grid=8)
    lev = ocfd.Level(boxes, g, eb=sp)
    og = gu.to_ogeom(g)
    p = ocfd.default_parm()
    fa = orc.new_level_fabs(boxes)
    orc.init_eb_inflow(boxes, fa, _valid_flags(boxes, g, sp), orc.parm())
    rng = np.random.default_rng(11 + visc + 2 * grav)
    for (lo, hi), f in zip(boxes, fa):
        v = f.valid(lo, hi)
        pert = 1.0 + 0.05 * rng.uniform(-1, 1, v[0].shape)
        v[0] *= pert
        v[2] = 0.1 * v[1] * rng.uniform(-1, 1, v[0].shape)
        v[4] *= 1.0 + 0.02 * rng.uniform(-1, 1, v[0].shape)
    assert orc.fill_ghost(boxes, fa, og, 0.0) == 0
    S = lev.new_mf()
    gu.upload(lev, S, fa)
    D = lev.new_mf(ngrow=0)
    flags = (ocfd.VISC if visc else 0) | (ocfd.GRAVITY if grav else 0) | (ocfd.EB_REDIST if redist else 0)
    lev.compute_dsdt(S, D, p, flags)
    got = gu.download(lev, D, ngrow=0)
    ebs = _eb_fabs(orc, boxes, g, sp)
    ref = []
    for (lo, hi), f, e in zip(boxes, fa, ebs):
        ref.append(orc.compute_dsdt_box_eb(lo, hi, f, orc.parm(), [1 / g.dx[d] for d in range(3)], e,
                                           visc, grav, redist))
    kinds = lev.box_types()
    for kd in (0, 1, 2):
        sel = [b for b in range(len(boxes)) if kinds[b] == kd]
        assert sel
        if kd == 2:
            assert all(np.abs(got[b]).max() == 0.0 for b in sel)
            continue
        err = gu.level_max_rel([got[b] for b in sel], [ref[b] for b in sel])
        print("box kind", kd, "max rel err", err)
        assert err < TOL64, (kd, err)


def _channel_run(orc, flags, visc, nsteps, nx=128, ny=32, max_grid=16, scheme=2):
    import ocfd
    from ocfd import cases
    boxes, g, kind, sp = cases.eb_channel(nx, ny, 2, max_grid=max_grid)
    lev = ocfd.Level(boxes, g, eb=sp)
    p = ocfd.default_parm()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    lev.init_state(kind, A, p)
    st_g, t_g, _ = ocfd.run(lev, A, B, C, p, 0.3, scheme, flags, nsteps=nsteps)
    got = gu.download(lev, A)
    og = gu.to_ogeom(g)
    fa = orc.new_level_fabs(boxes); fb = orc.new_level_fabs(boxes); fc = orc.new_level_fabs(boxes)
    orc.init_eb_inflow(boxes, fa, _valid_flags(boxes, g, sp), orc.parm())
    orc.set_eb(_eb_fabs(orc, boxes, g, sp), 1 if flags & ocfd.EB_REDIST else 0)
    try:
        st_o, t_o, _ = orc.run(boxes, fa, fb, fc, og, orc.parm(), 0.3, scheme, visc, 0, -1.0, nsteps)
    finally:
        orc.set_eb(None)
    errs = gu.level_rel_err([x[:, 4:-4, 4:-4, 4:-4] for x in got], [f.a[:, 4:-4, 4:-4, 4:-4] for f in fa])
    return (st_g, t_g), (st_o, t_o), errs, got


@pytest.mark.parametrize("visc", [0, 1])
def test_channel_steps_fp64(torch_cuda, orc, visc):
    """20 SSP-RK3 steps of the Mach 3 channel past the cylinder (with the
    small-cell scaling of do_redistribute)."""
    import ocfd
    flags = (ocfd.VISC if visc else 0) | ocfd.EB_REDIST
    (sg, tg), (so, to), errs, _ = _channel_run(orc, flags, visc, 20)
    worst = max(max(e) for e in errs)
    print("channel visc", visc, "(L1, Linf) per variable", errs)
    assert sg == so == 20
    assert abs(tg - to) <= 1e-12 * to
    assert worst < TOL64, worst


def test_channel_planar_equals_full_and_fp32(torch_cuda, orc):
    """OCFD_PLANAR on the z-uniform channel gives the same state bit for bit;
    the FP32-flux form (regular boxes FP32, cut boxes FP64) stays within 1e-4."""
    import ocfd
    R = ocfd.EB_REDIST
    _, _, _, full = _channel_run(orc, R, 0, 8)
    _, _, _, planar = _channel_run(orc, R | ocfd.PLANAR, 0, 8)
    for a, b in zip(full, planar):
        assert np.array_equal(a[:, 4:-4, 4:-4, 4:-4], b[:, 4:-4, 4:-4, 4:-4])
    _, _, errs32, _ = _channel_run(orc, R | ocfd.FLUX_FP32 | ocfd.PLANAR, 0, 8)
    print("channel fp32 (L1, Linf) per variable", errs32)
    assert errs32[0][0] < 1e-4
