"""GPU parity of the two-level AMR path against the AMR oracle: tagging, the
coarse part of the fine ghost fill, regrid data transfer, average-down, the
flux-register contents, and whole subcycled two-level runs (double Mach
reflection with regridding; a periodic viscous shear layer whose totals the
reflux keeps conserved).  FP64 tolerance 1e-10 relative; tags and box lists
exact."""
import ctypes

import numpy as np
import pytest

import amr_oracle as ao
import gpu_util as gu

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _dmr(orc, nx=64, ny=16, steps=4):
    """A coarse DMR level advanced a few single-level steps by the oracle."""
    from ocfd import cases
    boxes, g, kind = cases.dmr(nx, ny, 2, max_grid=(16, 8, 2))
    og = gu.to_ogeom(g)
    A = orc.new_level_fabs(boxes)
    B = orc.new_level_fabs(boxes)
    C = orc.new_level_fabs(boxes)
    orc.init_case(kind, boxes, A, og, orc.parm())
    orc.run(boxes, A, B, C, og, orc.parm(), 0.5, 2, 0, 0, -1.0, steps)
    orc.fill_ghost(boxes, A, og, 0.0)
    return boxes, g, og, kind, A


def test_tags_match_oracle(torch_cuda, orc):
    import ocfd
    from ocfd import amr
    boxes, g, og, kind, A = _dmr(orc)
    lev = ocfd.Level(boxes, g)
    S = lev.new_mf()
    gu.upload(lev, S, A)
    for thr in (0.05, 0.2, 1.0):
        got = amr.dense_tags(lev, amr.tag_dengrad(lev, S, thr).cpu().numpy())
        want = ao.tag_dengrad(boxes, A, og, thr)
        assert np.array_equal(got, want), thr
        assert 0 < want.sum() < want.size


def _fine_setup(orc, fboxes_c, boxes, g, og, A, seed=4):
    """Fine level over refined coarse boxes: the interpolated coarse state with
    a 1% random perturbation in the valid cells, random smooth ghosts."""
    from ocfd import amr
    gf = amr.refine_geom(g)
    ogf = gu.to_ogeom(gf)
    fb = ao.refine_boxes(fboxes_c)
    ff = orc.new_level_fabs(fb)
    rng = np.random.default_rng(seed)
    for f in ff:
        gu.random_state(f, rng)
    F = ao.interp_dense(ao.dense_grown(boxes, A, og), og)
    ao.interp_valid(fb, ff, ogf, F)
    for (lo, hi), f in zip(fb, ff):
        v = f.valid(lo, hi)
        v *= 1.0 + 0.01 * rng.uniform(-1, 1, v.shape)
    return gf, ogf, fb, ff


@pytest.mark.parametrize("case", ["dmr", "periodic"])
def test_fill_from_coarse_matches_oracle(torch_cuda, orc, case):
    import ocfd
    from ocfd import amr, cases
    if case == "dmr":
        boxes, g, og, kind, A = _dmr(orc)
        fine_c = [((4, 0, 0), (13, 5, 1)), ((14, 2, 0), (21, 9, 1)), ((30, 6, 0), (41, 15, 1))]
    else:
        boxes, g, kind = cases.kh3d(12, max_grid=(6, 6, 6))
        og = gu.to_ogeom(g)
        A = orc.new_level_fabs(boxes)
        orc.init_case(kind, boxes, A, og, orc.parm(), 2)
        orc.fill_ghost(boxes, A, og, 0.0)
        fine_c = [((0, 2, 0), (3, 7, 4)), ((8, 0, 6), (11, 5, 11)), ((4, 9, 3), (9, 11, 8))]
    gf, ogf, fb, ff = _fine_setup(orc, fine_c, boxes, g, og, A)
    lc, lf = ocfd.Level(boxes, g), ocfd.Level(fb, gf)
    Sc, Sf = lc.new_mf(), lf.new_mf()
    gu.upload(lc, Sc, A)
    gu.upload(lf, Sf, ff)
    t = 0.01
    lc.fill_boundary(Sc, t)
    lf.fill_boundary(Sf, t)
    amr.FluxRegister(lc, lf).fill_fine(Sf, Sc)
    got = gu.download(lf, Sf)
    orc.fill_ghost(boxes, A, og, t)
    orc.fill_ghost(fb, ff, ogf, t)
    ao.fill_from_coarse(fb, ff, ogf, ao.interp_dense(ao.dense_grown(boxes, A, og), og))
    err = gu.level_max_rel(got, [f.a for f in ff])
    print(case, "fine ghost fill max rel err", err)
    assert err < 1e-13


def test_regrid_transfer_and_avg_down_match_oracle(torch_cuda, orc):
    import ocfd
    from ocfd import amr
    boxes, g, og, kind, A = _dmr(orc)
    old_c = [((4, 0, 0), (13, 5, 1)), ((14, 2, 0), (21, 9, 1))]
    new_c = [((2, 0, 0), (11, 7, 1)), ((12, 0, 0), (25, 5, 1)), ((30, 4, 0), (35, 9, 1))]
    gf, ogf, ob, of = _fine_setup(orc, old_c, boxes, g, og, A)
    nb = ao.refine_boxes(new_c)
    lc, lo_, ln = ocfd.Level(boxes, g), ocfd.Level(ob, gf), ocfd.Level(nb, gf)
    Sc, So, Sn = lc.new_mf(), lo_.new_mf(), ln.new_mf()
    gu.upload(lc, Sc, A)
    gu.upload(lo_, So, of)
    amr.copy_level(ln, Sn, lo_, So)
    amr.FluxRegister(lc, ln).interp_fine(Sn, Sc, lo_)
    got = gu.download(ln, Sn)
    nf = orc.new_level_fabs(nb)
    skip = ao.copy_level(nb, nf, ob, of, ogf)
    ao.interp_valid(nb, nf, ogf, ao.interp_dense(ao.dense_grown(boxes, A, og), og), skip)
    v = lambda arrs: [a[:, 4:-4, 4:-4, 4:-4] for a in arrs]
    err = gu.level_max_rel(v(got), v([f.a for f in nf]))
    print("regrid transfer max rel err", err)
    assert err < 1e-13
    # average-down of the new fine level into the coarse level
    cf = amr.FluxRegister(lc, ln)
    cf.avg_down(Sn, Sc)
    gotc = gu.download(lc, Sc)
    ao.avg_down(nb, nf, boxes, A, og)
    errc = gu.level_max_rel(v(gotc), v([f.a for f in A]))
    print("avg_down max rel err", errc)
    assert errc < 1e-14


@pytest.mark.parametrize("visc,f32", [(0, 0), (1, 0), (1, 1)])
def test_flux_registers_match_oracle(torch_cuda, orc, visc, f32):
    """One coarse and one fine stage, each followed by a register add: the
    registers hold w*F of the oracle's face fluxes at every interface face."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.kh3d(12, max_grid=(6, 6, 12))
    og = gu.to_ogeom(g)
    A = orc.new_level_fabs(boxes)
    orc.init_case(kind, boxes, A, og, orc.parm(), 1)
    orc.fill_ghost(boxes, A, og, 0.0)
    fine_c = [((0, 2, 0), (3, 7, 4)), ((8, 0, 6), (11, 5, 11)), ((4, 9, 3), (9, 11, 8))]
    gf, ogf, fb, ff = _fine_setup(orc, fine_c, boxes, g, og, A)
    orc.fill_ghost(fb, ff, ogf, 0.0)
    ao.fill_from_coarse(fb, ff, ogf, ao.interp_dense(ao.dense_grown(boxes, A, og), og))
    lc, lf = ocfd.Level(boxes, g), ocfd.Level(fb, gf)
    Sc, Sf, Bc, Bf = lc.new_mf(), lf.new_mf(), lc.new_mf(), lf.new_mf()
    gu.upload(lc, Sc, A)
    gu.upload(lf, Sf, ff)
    p = ocfd.default_parm()
    flags = (ocfd.VISC if visc else 0) | (ocfd.FLUX_FP32 if f32 else 0)
    cf = amr.FluxRegister(lc, lf)
    assert cf.nfaces > 0
    cf.reset()
    wc, wf = 0.37, 0.11
    lc.rk_stage(Sc, None, Bc, 1.0, 0.0, 1e-3, p, flags)
    cf.add(0, wc, p, flags)
    lf.rk_stage(Sf, None, Bf, 1.0, 0.0, 1e-3, p, flags)
    cf.add(1, wf, p, flags)
    rc, rf = cf.registers()
    dxc = [1.0 / og.dx[d] for d in range(3)]
    dxf = [1.0 / ogf.dx[d] for d in range(3)]
    Fc = ao.dense_faces(boxes, [orc.compute_dsdt_box(lo, hi, f, orc.parm(), dxc, visc, 0, faces=True)[1]
                                for (lo, hi), f in zip(boxes, A)], og)
    Ff = ao.dense_faces(fb, [orc.compute_dsdt_box(lo, hi, f, orc.parm(), dxf, visc, 0, faces=True)[1]
                             for (lo, hi), f in zip(fb, ff)], ogf)
    nf = [ogf.dom_hi[d] + 1 for d in range(3)]
    want_c, want_f = np.zeros_like(rc), np.zeros_like(rf)
    for q, (b, i, j, k, d, up) in enumerate(cf.faces()):
        c = [i, j, k]
        c[d] += up
        want_c[q] = wc * Fc[d][:, c[2], c[1], c[0]]
        t1, t2 = (1 if d == 0 else 0), (1 if d == 2 else 2)
        for s, (a1, a2) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
            gcell = [2 * i, 2 * j, 2 * k]
            gcell[d] = 2 * (i, j, k)[d] + 2 if up else 2 * (i, j, k)[d] - 1
            gcell[t1] += a1
            gcell[t2] += a2
            gcell[d] %= nf[d]
            face = list(gcell)
            if not up:
                face[d] += 1
            want_f[q, s] = wf * Ff[d][:, face[2], face[1], face[0]]
    tol = 1e-4 if f32 else 1e-11
    scale = max(np.abs(want_c).max(), np.abs(want_f).max())
    ec = np.abs(rc - want_c).max() / scale
    ef = np.abs(rf - want_f).max() / scale
    print("registers visc", visc, "fp32", f32, "max rel err", ec, ef)
    assert ec < tol and ef < tol


@pytest.mark.parametrize("flags", [0, 1, 5])   # inviscid, viscous, viscous FP32-flux
def test_compute_dsdt_feeds_flux_registers(torch_cuda, orc, flags):
    """compute_dSdt with its register arguments through the C ABI: the dSdt
    equals the plain compute, and the registers equal an explicit add after
    each compute (bit for bit), both sides of one interface, two stages each;
    a register on the wrong side of the level is refused."""
    import torch
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.kh3d(12, max_grid=(6, 6, 12))
    og = gu.to_ogeom(g)
    A = orc.new_level_fabs(boxes)
    orc.init_case(kind, boxes, A, og, orc.parm(), 1)
    orc.fill_ghost(boxes, A, og, 0.0)
    fine_c = [((0, 2, 0), (3, 7, 4)), ((8, 0, 6), (11, 5, 11))]
    gf, ogf, fb, ff = _fine_setup(orc, fine_c, boxes, g, og, A)
    orc.fill_ghost(fb, ff, ogf, 0.0)
    ao.fill_from_coarse(fb, ff, ogf, ao.interp_dense(ao.dense_grown(boxes, A, og), og))
    lc, lf = ocfd.Level(boxes, g), ocfd.Level(fb, gf)
    Sc, Sf = lc.new_mf(), lf.new_mf()
    gu.upload(lc, Sc, A)
    gu.upload(lf, Sf, ff)
    assert ocfd.VISC == 1 and ocfd.FLUX_FP32 == 4
    p = ocfd.default_parm()
    cf1, cf2 = amr.FluxRegister(lc, lf), amr.FluxRegister(lc, lf)
    assert cf1.nfaces > 0
    cf1.reset()
    cf2.reset()
    Dc1, Dc2, Df1, Df2 = lc.new_mf(ngrow=0), lc.new_mf(ngrow=0), lf.new_mf(ngrow=0), lf.new_mf(ngrow=0)
    for w in (0.25, 0.5):
        lc.compute_dsdt_fr(Sc, Dc1, p, flags, dt=w, fine=cf1)
        lf.compute_dsdt_fr(Sf, Df1, p, flags, dt=0.5 * w, current=cf1)
        lc.compute_dsdt(Sc, Dc2, p, flags)
        cf2.add(0, w, p, flags)
        lf.compute_dsdt(Sf, Df2, p, flags)
        cf2.add(1, 0.5 * w, p, flags)
    torch.cuda.synchronize()
    assert torch.equal(Dc1, Dc2) and torch.equal(Df1, Df2)
    (rc1, rf1), (rc2, rf2) = cf1.registers(), cf2.registers()
    assert np.array_equal(rc1, rc2) and np.array_equal(rf1, rf2)
    assert np.abs(rc1).max() > 0 and np.abs(rf1).max() > 0
    # wrong side: the fine level is not the coarse side of its own register
    with pytest.raises(RuntimeError, match="status 1"):
        lf.compute_dsdt_fr(Sf, Df1, p, flags, dt=1.0, fine=cf1)
    with pytest.raises(RuntimeError, match="status 1"):
        lc.compute_dsdt_fr(Sc, Dc1, p, flags, dt=1.0, current=cf1)


@pytest.mark.parametrize("periodic_x", [0, 1])
def test_tag_nest_flags_match_host(torch_cuda, orc, periodic_x):
    """The device flags of a partial level: bit 0 equals ocfd_tag_dengrad's
    tag, bit 1 the host's dense proper nesting region (n_proper 1 and 2) at
    the level's cells."""
    import torch
    import ocfd
    from ocfd import amr, cases
    boxes, g, og, kind, A = _dmr(orc)
    g.periodic[0] = periodic_x
    part = [b for q, b in enumerate(boxes) if q % 3 != 1]
    keep = [q for q, b in enumerate(boxes) if q % 3 != 1]
    lev = ocfd.Level(part, g)
    S = lev.new_mf()
    gu.upload(lev, S, [A[q] for q in keep])
    tags = amr.tag_dengrad(lev, S, 0.05).cpu().numpy()
    for n_proper in (1, 2):
        flags = amr.tag_dengrad_nest(lev, S, 0.05, n_proper).cpu().numpy()
        assert np.array_equal(flags & 1, tags)
        pn = amr.proper_nesting(part, g, n_proper).astype(np.uint8)
        want = np.concatenate([pn[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].reshape(-1) for lo, hi in part])
        assert np.array_equal(flags >> 1, want), n_proper
        assert 0 < want.sum() < want.size
    assert tags.sum() > 0
    torch.cuda.synchronize()


def _compare_runs(T, O, tol):
    key = lambda b: (tuple(b[0]), tuple(b[1]))
    assert sorted(map(key, T.fine.boxes)) == sorted(map(key, O.fboxes))
    v = lambda arrs: [a[:, 4:-4, 4:-4, 4:-4] for a in arrs]
    ref_f = dict(zip(map(key, O.fboxes), (f.a for f in O.S[1])))
    ec = gu.level_rel_err(v(gu.download(T.coarse, T.S[0])), v([f.a for f in O.S[0]]))
    ef = gu.level_rel_err(v(gu.download(T.fine, T.S[1])), v([ref_f[key(b)] for b in T.fine.boxes]))
    worst = max(max(max(e) for e in ec), max(max(e) for e in ef))
    print("coarse (L1, Linf)", ec)
    print("fine (L1, Linf)", ef)
    assert worst < tol, worst


@pytest.mark.parametrize("scheme", [2, 1])
def test_two_level_dmr_matches_oracle(torch_cuda, orc, scheme):
    """Double Mach reflection, 64x16x2 coarse + regridded fine level
    (refine_dengrad 0.05, regrid_int 2, blocking_factor 2, max_grid_size 32),
    6 coarse steps: grids identical, states within 1e-10."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    og = gu.to_ogeom(g)
    kw = dict(scheme=scheme, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, max_grid_size=32)
    T = amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.5, kind, **kw)
    O = ao.TwoLevel(boxes, og, orc.parm(), 0.5, kind, **kw)
    for s in range(6):
        dt_g, dt_o = T.step(), O.step()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (s, dt_g, dt_o)
    assert abs(T.time - O.time) <= 1e-12 * O.time
    _compare_runs(T, O, TOL64)


def test_two_level_viscous_periodic_conserves_and_matches(torch_cuda, orc):
    """Viscous 3D shear layer, 12^3 coarse + fixed fine boxes touching the
    periodic faces: totals of mass, momentum and energy conserved to rounding
    on the device, and the states match the oracle."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.kh3d(12, max_grid=(6, 12, 12))
    og = gu.to_ogeom(g)
    fine = [((0, 1, 2), (3, 4, 4)), ((10, 1, 2), (11, 6, 5)), ((4, 8, 0), (7, 11, 11))]
    kw = dict(scheme=2, regrid_int=0, fine_boxes=fine, seed=3)
    T = amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.4, kind, flags=ocfd.VISC, **kw)
    O = ao.TwoLevel(boxes, og, orc.parm(), 0.4, kind, visc=1, **kw)

    def totals():
        tot = np.zeros(5)
        cov = np.zeros((12, 12, 12), dtype=bool)
        for lo, hi in fine:
            cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        for (lo, hi), a in zip(boxes, gu.download(T.coarse, T.S[0])):
            v = a[:, 4:-4, 4:-4, 4:-4]
            tot += 8.0 * v[:, ~cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]].sum(axis=1)
        for a in gu.download(T.fine, T.S[1]):
            tot += a[:, 4:-4, 4:-4, 4:-4].reshape(5, -1).sum(axis=1)
        return tot

    t0 = totals()
    for _ in range(4):
        T.step()
        O.step()
    t1 = totals()
    scale = np.array([t0[0], t0[0], t0[0], t0[0], t0[4]])
    print("relative drift of the totals", (t1 - t0) / scale)
    assert (np.abs(t1 - t0) / scale < 1e-13).all()
    _compare_runs(T, O, TOL64)


def test_two_level_fp32_flux_and_planar(torch_cuda, orc):
    """FP32-flux form of the two-level DMR within 1e-4 (L1) of the FP64 form;
    planar stepping of the z-uniform flow equal to the full form bit for bit."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    fine = [((0, 0, 0), (21, 9, 1)), ((22, 0, 0), (31, 5, 1))]
    runs = []
    for flags in (0, ocfd.FLUX_FP32, ocfd.PLANAR):
        T = amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.5, kind, flags=flags, regrid_int=0, fine_boxes=fine)
        for _ in range(4):
            T.step()
        runs.append([a[:, 4:-4, 4:-4, 4:-4] for a in gu.download(T.fine, T.S[1])])
    e32 = gu.level_rel_err(runs[1], runs[0])
    print("fp32 vs fp64 fine level (L1, Linf)", e32)
    assert e32[0][0] < 1e-4
    # z-uniform data: skipping the z sweep changes nothing (FP64 partial sums)
    for a, b in zip(runs[0], runs[2]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_planar_hierarchy_refuses_fine_grids_short_of_the_z_domain():
    """Planar stepping builds fine grids that span the domain in z, and refuses
    one that covers half of it (its register would hold z faces)."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    T = amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.5, kind, flags=ocfd.PLANAR, regrid_int=0,
                     fine_boxes=[((0, 0, 0), (21, 9, 1))])
    assert T.cfs[1].nfaces > 0
    del T
    with pytest.raises(ValueError, match="span the domain in z"):
        amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.5, kind, flags=ocfd.PLANAR, regrid_int=0,
                     fine_boxes=[((0, 0, 0), (21, 9, 0))])


@pytest.mark.parametrize("flags", ["PLANAR|VISC", "PLANAR|FLUX_FP32|VISC"])
def test_two_level_planar_fill_depth_viscous(torch_cuda, orc, flags):
    """The one-layer planar fill (first-layer sweeps, first-layer register
    faces whose viscous terms read the z neighbours the x tiles wrote) equals
    the same planar stepping with the four-layer fill, bit for bit, in the
    viscous FP64 and FP32-flux forms."""
    import ocfd
    from ocfd import amr, cases
    f = 0
    for name in flags.split("|"):
        f |= getattr(ocfd, name)
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    fine = [((0, 0, 0), (21, 9, 1)), ((22, 0, 0), (31, 5, 1))]
    runs = []
    for pf in (True, False):
        T = amr.TwoLevel(boxes, g, ocfd.default_parm(), 0.5, kind, flags=f, regrid_int=0, fine_boxes=fine,
                         planar_fill=pf)
        for _ in range(4):
            T.step()
        runs.append([a[:, 4:-4, 4:-4, 4:-4] for lev, S in ((T.levels[0], T.S[0]), (T.fine, T.S[1]))
                     for a in gu.download(lev, S)])
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def _compare_hier(T, O, tol):
    key = lambda b: (tuple(b[0]), tuple(b[1]))
    assert T.finest == O.finest
    v = lambda arrs: [a[:, 4:-4, 4:-4, 4:-4] for a in arrs]
    worst = 0.0
    for l in range(T.finest + 1):
        assert sorted(map(key, T.levels[l].boxes)) == sorted(map(key, O.boxes[l])), l
        ref = dict(zip(map(key, O.boxes[l]), (f.a for f in O.S[l])))
        e = gu.level_rel_err(v(gu.download(T.levels[l], T.S[l])), v([ref[key(b)] for b in T.levels[l].boxes]))
        print("level", l, "(L1, Linf)", e)
        worst = max(worst, max(max(x) for x in e))
    assert worst < tol, worst


def test_three_level_dmr_matches_oracle(torch_cuda, orc):
    """Double Mach reflection with two AMR levels (64x16x2 base, regrid every 2
    steps of each level, max_grid_size 32): grids of every level identical,
    states within 1e-10 after 4 coarse steps."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    kw = dict(scheme=2, max_level=2, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, max_grid_size=32)
    T = amr.Hierarchy(boxes, g, ocfd.default_parm(), 0.5, kind, **kw)
    O = ao.Hierarchy(boxes, gu.to_ogeom(g), orc.parm(), 0.5, kind, **kw)
    for s in range(4):
        dt_g, dt_o = T.step(), O.step()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (s, dt_g, dt_o)
    assert T.finest == 2
    _compare_hier(T, O, TOL64)


def test_three_level_dmr_uneven_planar_matches_oracle(torch_cuda, orc):
    """The bench's hierarchy at a quarter of its base: DMR 256x64x2 in seeded
    uneven boxes of 16-64 cells, max_level 2, regrid every 2 steps of each
    level, blocking factor 2, max_grid_size 64, planar stepping (one-layer
    ghost fills, first-layer sweeps) against the oracle's full 3D run: grids
    of every level identical, states within 1e-10 after 4 coarse steps (two
    regrids of each level)."""
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(256, 64, 2, sizes=(16, 32, 64), seed=1234)
    kw = dict(scheme=2, max_level=2, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, grid_eff=0.7,
              max_grid_size=64)
    T = amr.Hierarchy(boxes, g, ocfd.default_parm(), 0.5, kind, flags=ocfd.PLANAR, **kw)
    O = ao.Hierarchy(boxes, gu.to_ogeom(g), orc.parm(), 0.5, kind, **kw)
    for s in range(4):
        dt_g, dt_o = T.step(), O.step()
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, (s, dt_g, dt_o)
    assert T.finest == 2 and len(T.levels[2].boxes) > 8
    _compare_hier(T, O, TOL64)


def test_three_level_viscous_periodic_conserves_and_matches(torch_cuda, orc):
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.kh3d(8, max_grid=(8, 8, 8))
    fine = [[((0, 1, 1), (7, 6, 5))], [((0, 4, 4), (3, 9, 7)), ((14, 4, 4), (15, 9, 7)), ((6, 3, 3), (9, 7, 5))]]
    kw = dict(scheme=2, max_level=2, regrid_int=0, fine_boxes=fine, seed=5)
    T = amr.Hierarchy(boxes, g, ocfd.default_parm(), 0.4, kind, flags=ocfd.VISC, **kw)
    O = ao.Hierarchy(boxes, gu.to_ogeom(g), orc.parm(), 0.4, kind, visc=1, **kw)

    def totals():
        tot = np.zeros(5)
        for l in range(T.finest + 1):
            geo = T.geoms[l]
            n = [geo.dom_hi[d] + 1 for d in range(3)]
            cov = np.zeros((n[2], n[1], n[0]), dtype=bool)
            if l < T.finest:
                for lo, hi in T.levels[l + 1].boxes:
                    cov[lo[2] // 2:hi[2] // 2 + 1, lo[1] // 2:hi[1] // 2 + 1, lo[0] // 2:hi[0] // 2 + 1] = True
            for (lo, hi), a in zip(T.levels[l].boxes, gu.download(T.levels[l], T.S[l])):
                vv = a[:, 4:-4, 4:-4, 4:-4]
                tot += 8.0 ** (T.finest - l) * vv[:, ~cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]].sum(1)
        return tot

    t0 = totals()
    for _ in range(2):
        T.step()
        O.step()
    t1 = totals()
    scale = np.array([t0[0]] * 4 + [t0[4]])
    print("relative drift of the totals", (t1 - t0) / scale)
    assert (np.abs(t1 - t0) / scale < 1e-13).all()
    _compare_hier(T, O, TOL64)


def test_hierarchy_one_call_forms_equal_per_level_calls(torch_cuda, orc):
    """The hierarchy's batched forms against the per-level calls they stand
    for, bit for bit, on the uneven planar DMR after two coarse steps (both
    levels regridded): the CFL reductions of every level in one call
    (ocfd_amr_estdt) against ocfd_estdt level by level; the finest level's
    ghost fill at the current time (level and coarse/fine fill in one launch
    over its ghost lists) against ocfd_fill_boundary followed by
    ocfd_cf_fill_fine; and every register's faces listed in (coarse box, k, j,
    i, direction, side) order, the order of a global sort."""
    import torch
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(256, 64, 2, sizes=(16, 32, 64), seed=1234)
    kw = dict(scheme=2, max_level=2, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, grid_eff=0.7,
              max_grid_size=64)
    T = amr.Hierarchy(boxes, g, ocfd.default_parm(), 0.5, kind, flags=ocfd.PLANAR, **kw)
    for _ in range(2):
        T.step()
    torch.cuda.synchronize()
    n = T.finest + 1
    assert n == 3
    lv = (ctypes.c_void_p * n)(*[T.levels[l].handle for l in range(n)])
    st = (ctypes.c_void_p * n)(*[T.S[l].data_ptr() for l in range(n)])
    out = (ctypes.c_double * n)()
    ocfd.check(ocfd.lib().ocfd_amr_estdt(n, lv, st, 0.5, ctypes.byref(T.parm), out, None), "ocfd_amr_estdt")
    per = [T.levels[l].estdt(T.S[l], 0.5, T.parm) for l in range(n)]
    assert list(out) == per, (list(out), per)

    l, t = T.finest, T.time
    X1 = T.S[l].clone()
    T.fill(l, X1, t)
    X2 = T.S[l].clone()
    T.fill(l - 1, None, t)
    T.levels[l].fill_boundary(X2, t)
    T.cfs[l].fill_fine(X2, T.S[l - 1])
    torch.cuda.synchronize()
    assert torch.equal(X1, X2)

    for m in range(1, n):
        f = T.cfs[m].faces()
        assert len(f) > 0
        keys = [(b, k, j, i, d, up) for b, i, j, k, d, up in f.tolist()]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
