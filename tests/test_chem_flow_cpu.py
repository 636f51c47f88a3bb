"""Pins of the oracle's multispecies (CHEM) flow path, EBR::compute_dSdt_multi
(Source/CHEM/FlowAdvance.cpp:92-496) with c2prim_rgas (Kernels.H:42-81),
the species Steger-Warming split, WENO-Z5 for the 5 + 9 components,
Fickian diffusion and the multispecies viscous flux (ChemViscous.H).

The transport functions are pinned bitwise against the reference build in
test_chem_oracle.py; here the assembled box path is pinned by exact
conservation on a periodic box, free-stream preservation, the single-species
identity between the species and mass fluxes, the viscous heating of a
Couette flow with the mixture viscosity, and the direction of diffusion."""
import ctypes

import numpy as np
import pytest

import ocfd

RU = 8.31446261815324


def _og(orc, g):
    og = orc.OGeom()
    ctypes.memmove(ctypes.byref(og), ctypes.byref(g), ctypes.sizeof(og))
    return og


def _state(orc, n, Y, T, p, vel):
    """Conserved state and partial densities on an (n,n,n) grid from Y (9,...),
    T, p and velocity fields."""
    mw = orc.chem_tables()[0]
    rho = p / (RU * T * (Y / mw[:, None, None, None]).sum(0))
    rhoi = rho * Y
    e = np.array([orc.thermo("CKUBMS", T.flat[c], rhoi.reshape(9, -1)[:, c]) for c in range(T.size)]).reshape(T.shape)
    u, v, w = vel
    U = np.stack([rho, rho * u, rho * v, rho * w, e + 0.5 * rho * (u * u + v * v + w * w)])
    return U, rhoi


def _box(orc, n, U, rhoi, periodic=(1, 1, 1), L=0.01):
    g = ocfd.make_geom((n, n, n), (0, 0, 0), (L, L, L), periodic=periodic)
    og = _og(orc, g)
    boxes = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    A = orc.new_level_fabs(boxes)
    As = orc.new_spec_fabs(boxes)
    lo, hi = boxes[0]
    A[0].valid(lo, hi)[:] = U
    As[0].valid(lo, hi)[:] = rhoi
    assert orc.fill_ghost(boxes, A, og, 0.0) == 0
    assert orc.fill_ghost_spec(boxes, As, og, 0.0) == 0
    return g, boxes, A, As


def _grid(n):
    return np.meshgrid(*(np.arange(n),) * 3, indexing="ij")


@pytest.mark.parametrize("visc", [0, 1])
def test_periodic_box_conserves_every_component(orc, visc):
    n = 16
    z, y, x = _grid(n)
    Y = np.zeros((9, n, n, n))
    Y[0] = 0.02 + 0.01 * np.sin(2 * np.pi * x / n)
    Y[1] = 0.22
    Y[5] = 1e-3 * (1 + np.cos(2 * np.pi * z / n))
    Y[8] = 1 - Y.sum(0)
    T = 900 + 100 * np.cos(2 * np.pi * y / n)
    p = 101325 * (1 + 0.1 * np.sin(2 * np.pi * z / n))
    vel = (10 * np.sin(2 * np.pi * y / n), 5 * np.cos(2 * np.pi * x / n), 3 * np.sin(2 * np.pi * x / n))
    U, rhoi = _state(orc, n, Y, T, p, vel)
    g, boxes, A, As = _box(orc, n, U, rhoi)
    D, Ds = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, visc)
    for arr in (D, Ds):
        tot = arr.reshape(arr.shape[0], -1).sum(1)
        mag = np.abs(arr).reshape(arr.shape[0], -1).sum(1)
        assert np.all(np.abs(tot) <= 1e-12 * np.maximum(mag, 1e-300))
    assert np.abs(D).max() > 0


def test_free_stream(orc):
    n = 8
    Y = np.zeros((9, n, n, n))
    Y[0], Y[1], Y[8] = 0.03, 0.23, 0.74
    T = np.full((n, n, n), 1200.0)
    p = np.full((n, n, n), 2 * 101325.0)
    vel = (np.full((n, n, n), 120.0), np.full((n, n, n), -40.0), np.full((n, n, n), 7.0))
    U, rhoi = _state(orc, n, Y, T, p, vel)
    g, boxes, A, As = _box(orc, n, U, rhoi)
    for visc in (0, 1):
        D, Ds = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, visc)
        assert np.abs(D).max() == 0.0 and np.abs(Ds).max() == 0.0


def test_single_species_flux_equals_mass_flux(orc):
    """Pure N2: the species split of rho_N2 = rho is the mass split, so the
    N2 divergence equals the density divergence bit for bit (inviscid)."""
    n = 12
    z, y, x = _grid(n)
    Y = np.zeros((9, n, n, n))
    Y[8] = 1.0
    T = 800 + 200 * np.sin(2 * np.pi * x / n) * np.cos(2 * np.pi * y / n)
    p = 101325 * (1 + 0.2 * np.cos(2 * np.pi * z / n))
    vel = (50 * np.sin(2 * np.pi * z / n), 30 * np.cos(2 * np.pi * x / n), 20 * np.sin(2 * np.pi * y / n))
    U, rhoi = _state(orc, n, Y, T, p, vel)
    rhoi[8] = U[0]
    g, boxes, A, As = _box(orc, n, U, rhoi)
    D, Ds = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, 0)
    assert np.array_equal(Ds[8], D[0])
    assert np.abs(Ds[:8]).max() == 0.0


def test_couette_heating_with_mixture_viscosity(orc):
    """u = a y at uniform T, p and composition, given on the grown box
    (ghosts included): momentum dSdt = 0 and energy dSdt = mu_mix a^2 with
    mu_mix from mixtureProperties."""
    n = 12
    z, y, x = _grid(n)
    Y = np.zeros((9, n, n, n))
    Y[0], Y[1], Y[8] = 0.02, 0.23, 0.75
    T = np.full((n, n, n), 600.0)
    p = np.full((n, n, n), 101325.0)
    dx = 0.01 / n
    a = 2.0e4
    yc = (np.arange(-4, n + 4) + 0.5) * dx
    # state on the grown box directly: linear u in y, including ghosts
    ng = n + 8
    zz, yy, xx = np.meshgrid(*(np.arange(ng),) * 3, indexing="ij")
    Yg = np.zeros((9, ng, ng, ng))
    Yg[0], Yg[1], Yg[8] = 0.02, 0.23, 0.75
    Ug, rg = _state(orc, ng, Yg, np.full((ng,) * 3, 600.0), np.full((ng,) * 3, 101325.0),
                    (a * yc[yy], np.zeros((ng,) * 3), np.zeros((ng,) * 3)))
    g = ocfd.make_geom((n, n, n), (0, 0, 0), (0.01, 0.01, 0.01), periodic=(1, 1, 1))
    boxes = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    A = orc.new_level_fabs(boxes)
    As = orc.new_spec_fabs(boxes)
    A[0].a[:] = Ug
    As[0].a[:] = rg
    D, Ds = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / dx] * 3, 1)
    X = orc.CKYTX(Yg[:, 0, 0, 0])
    mu, _ = orc.mixture_properties(600.0, X)
    rel = np.abs(D[4] - mu * a * a).max() / (mu * a * a)
    assert rel < 1e-7, rel          # T comes from a Newton solve stopped at 1e-7 K
    assert np.abs(D[1:4]).max() < 1e-9 * mu * a * a / (a * 0.01)
    assert np.abs(Ds).max() < 1e-9 * np.abs(rg).max() / 1e-3


def test_diffusion_flows_down_the_gradient(orc):
    """A Gaussian H2 bump at rest, uniform T and p: the viscous minus the
    inviscid species divergence is Fickian, negative at the peak."""
    n = 16
    z, y, x = _grid(n)
    Y = np.zeros((9, n, n, n))
    Y[0] = 0.01 + 0.02 * np.exp(-((x - n / 2 + 0.5) ** 2) / 8.0)
    Y[1] = 0.23
    Y[8] = 1 - Y[0] - Y[1]
    T = np.full((n, n, n), 800.0)
    p = np.full((n, n, n), 101325.0)
    zero = np.zeros((n, n, n))
    U, rhoi = _state(orc, n, Y, T, p, (zero, zero, zero))
    g, boxes, A, As = _box(orc, n, U, rhoi)
    _, Ds0 = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, 0)
    _, Ds1 = orc.compute_dsdt_multi_box(*boxes[0], A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, 1)
    dH2 = (Ds1 - Ds0)[0]
    c = n // 2
    assert dH2[c, c, c - 1] < 0 and dH2[c, c, c] < 0
    assert dH2[c, c, 0] > 0


# ---------------------------------------------------------------------------
# pins of the EB-CHEM species split, the Fickian flux and the diffusive
# enthalpy flux by facts independent of their formulas
# ---------------------------------------------------------------------------
def test_species_split_single_species_equals_van_leer_density_flux(orc):
    """FlowAdvanceBox.cpp:174-223 splits each partial density as the Van Leer
    mass flux does: a mixture of one species gives that species the density
    flux of the Van Leer split (EBkernels.H:73-171, pinned by its
    consistency with the Euler flux) bit for bit, in every Mach regime, and
    the other species nothing; for a mixture the species fluxes sum to the
    density flux."""
    par = orc.parm()
    rng = np.random.default_rng(21)
    n = 400
    cL, cR = rng.uniform(300, 900, n), rng.uniform(300, 900, n)
    ul, ur = rng.uniform(-3, 3, n) * cL, rng.uniform(-3, 3, n) * cR
    rhoL, rhoR = rng.uniform(0.1, 5, n), rng.uniform(0.1, 5, n)
    ql = np.column_stack([rhoL, ul, rng.uniform(-50, 50, n), rng.uniform(-50, 50, n), np.full(n, 1e5),
                          np.full(n, 300.0), cL])
    qr = np.column_stack([rhoR, ur, rng.uniform(-50, 50, n), rng.uniform(-50, 50, n), np.full(n, 1e5),
                          np.full(n, 300.0), cR])
    mass = orc.eb_flux(0, ql, qr, par)[:, 0]
    regimes = {(bool(abs(a) < 1), bool(abs(b) < 1)) for a, b in zip(ul / cL, ur / cR)}
    assert len(regimes) == 4
    for k in range(9):
        rl = np.zeros((n, 9)); rr = np.zeros((n, 9))
        rl[:, k], rr[:, k] = rhoL, rhoR
        F = orc.vl_species_flux(ul, ur, cL, cR, rl, rr)
        assert np.array_equal(F[:, k], mass), k
        assert (np.delete(F, k, axis=1) == 0).all()
    Yl, Yr = rng.uniform(0, 1, (n, 9)), rng.uniform(0, 1, (n, 9))
    rl, rr = Yl / Yl.sum(1, keepdims=True) * rhoL[:, None], Yr / Yr.sum(1, keepdims=True) * rhoR[:, None]
    F = orc.vl_species_flux(ul, ur, cL, cR, rl, rr)
    assert (np.abs(F.sum(1) - mass) <= 1e-13 * (np.abs(mass) + np.abs(F).sum(1))).all()


@pytest.mark.parametrize("visc", [0, 1])
def test_ebchem_box_conserves_species_mass(orc, visc):
    """One periodic box around a cylinder through the EB-CHEM path: the face
    fluxes telescope with the area fractions and the species wall term
    carries no mass, so each species' total changes by rounding only."""
    n = 16
    L = 0.01
    rng = np.random.default_rng(5 + visc)
    Y = rng.uniform(0.05, 1.0, (9, n, n, n))
    Y /= Y.sum(0)
    T = rng.uniform(700, 900, (n, n, n))
    p = rng.uniform(0.9e5, 1.1e5, (n, n, n))
    vel = tuple(rng.uniform(-20, 20, (n, n, n)) for _ in range(3))
    U, rhoi = _state(orc, n, Y, T, p, vel)
    g, boxes, A, As = _box(orc, n, U, rhoi, L=L)
    sp = ocfd.eb_cylinder(0.5 * L, 0.5 * L, 0.22 * L)
    lo, hi = boxes[0]
    glo, ghi = orc.grown(lo, hi)
    gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
    eb = orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"])
    assert orc.eb_box_kind(lo, hi, eb) == 1
    D, Ds = orc.compute_dsdt_multi_box_eb(lo, hi, A[0], As[0], orc.parm(), [1 / g.dx[0]] * 3, eb, visc)
    assert (D == 0).all()                       # the reference never copies the flow part
    vf = eb.arrs[0][4:-4, 4:-4, 4:-4]
    tot = (Ds * vf[None]).reshape(9, -1).sum(1)
    mag = (np.abs(Ds) * vf[None]).reshape(9, -1).sum(1)
    assert (np.abs(tot) <= 1e-12 * mag).all(), tot / mag
    assert (mag > 0).all()


def _linear_composition_box(orc, n, eps, T=800.0, p=101325.0, L=0.01, dTdx=0.0):
    """Grown box (ghosts included) at rest, uniform p, with Y_H2 rising and
    Y_N2 falling linearly in x at the rate eps (1/m), T = T + dTdx x."""
    ng = n + 8
    dx = L / n
    xc = (np.arange(-4, n + 4) + 0.5) * dx
    Y = np.zeros((9, ng, ng, ng))
    Y[0], Y[1], Y[2] = 0.02, 0.2, 0.05
    Y[0] += eps * xc[None, None, :]
    Y[8] = 1.0 - Y[0] - Y[1] - Y[2]
    zero = np.zeros((ng,) * 3)
    Tg = T + dTdx * np.broadcast_to(xc[None, None, :], (ng, ng, ng))
    U, rg = _state(orc, ng, Y, Tg, np.full((ng,) * 3, p), (zero, zero, zero))
    boxes = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    A, As = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    A[0].a[:] = U
    As[0].a[:] = rg
    return boxes, A, As, dx, Y


def test_fickian_flux_is_ficks_law(orc):
    """J_k = -rho D_k dY_k/dx at every x face (ChemViscous.H:108-199), with
    D_k from getMixDiffCoeffsMass (pinned bitwise against the reference
    build) at the face's composition: a small linear H2/N2 gradient at rest,
    uniform T and p; species without a gradient carry no flux."""
    n, eps, T, p = 8, 0.02, 800.0, 101325.0
    boxes, A, As, dx, Y = _linear_composition_box(orc, n, eps, T, p)
    F, Fs, Fd = orc.compute_multi_faces(*boxes[0], A[0], As[0], orc.parm(), [1 / dx] * 3, 1)
    J = Fd[0]                                              # x faces, (9, n, n, n+1)
    mw = orc.chem_tables()[0]
    for i in range(n + 1):
        yf = Y[:, 4, 4, i + 3] * 0.5 + Y[:, 4, 4, i + 4] * 0.5   # composition at the face
        rho = p / (RU * T * (yf / mw).sum())
        Dk = orc.mix_diff_coeffs(T, p, orc.CKYTX(yf))
        for k, slope in ((0, eps), (8, -eps)):
            want = -rho * Dk[k] * slope
            assert abs(J[k, 0, 0, i] / want - 1) < 1e-7, (i, k, J[k, 0, 0, i], want)
        others = [1, 2, 3, 4, 5, 6, 7]
        assert np.abs(J[others, :, :, i]).max() <= 1e-9 * abs(J[0, 0, 0, i])
    # no gradient along y and z
    assert np.abs(Fd[1]).max() <= 1e-9 * np.abs(J).max() and np.abs(Fd[2]).max() <= 1e-9 * np.abs(J).max()


def test_species_enthalpy_is_internal_energy_plus_RT(orc):
    """speciesEnthalpy (LiDryer.H:616-695) returns h/R: for the same NASA7
    fits h/R = T (e/(RT) + 1), with e/(RT) from the reference build's own
    speciesInternalEnergy, in both temperature branches."""
    ref = orc.ref_chem()
    if ref is None:
        pytest.skip("oracle/_ref/libchemref.so not built (needs the reference sources)")
    f = ref._Z21speciesInternalEnergyPdS_
    dp = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [dp, dp]
    rng = np.random.default_rng(3)
    for T in np.concatenate([rng.uniform(250, 999, 40), rng.uniform(1001, 3500, 40)]):
        e = np.zeros(9)
        tc = np.array([0.0, T, T * T, T * T * T, T * T * T * T])
        f(e.ctypes.data_as(dp), tc.ctypes.data_as(dp))
        h = np.zeros(9)
        orc.thermo("species_enthalpy", T, h)
        assert np.allclose(h, T * (e + 1.0), rtol=1e-13, atol=0), T


def test_diffusive_enthalpy_flux_in_the_energy_flux(orc):
    """compute_visc_x_multi (ChemViscous.H:201-290) at rest: the viscous part
    of the energy flux is -lambda dT/dx - sum_k h_k(T_f) J_k with h_k the
    species enthalpy in the reference's h/R units at the face temperature
    T_f (mean of the two cells) and J_k the Fickian flux of the same face;
    lambda from mixtureProperties (pinned bitwise against the reference
    build), T of each cell from the pinned Newton solve; a temperature
    gradient of 2.5 K per cell tells T_f from either cell's T."""
    n, eps, T, p = 8, 2.0, 800.0, 101325.0
    boxes, A, As, dx, Y = _linear_composition_box(orc, n, eps, T, p, dTdx=2000.0)
    par = orc.parm()
    F1, _, Fd = orc.compute_multi_faces(*boxes[0], A[0], As[0], par, [1 / dx] * 3, 1)
    F0, _, _ = orc.compute_multi_faces(*boxes[0], A[0], As[0], par, [1 / dx] * 3, 0)
    a = A[0].a
    s = As[0].a

    def cell_T(i):
        rhoi = s[:, 4, 4, i]
        ke = 0.5 * (a[1, 4, 4, i] ** 2 + a[2, 4, 4, i] ** 2 + a[3, 4, 4, i] ** 2) / a[0, 4, 4, i]
        return orc.T_given_ey(a[4, 4, 4, i] - ke, rhoi)[0]

    checked = 0
    for i in range(n + 1):
        gl, gr = i + 3, i + 4
        Tl, Tr = cell_T(gl), cell_T(gr)
        lam = [orc.mixture_properties(t, orc.CKYTX(s[:, 4, 4, g] / s[:, 4, 4, g].sum()))[1] for t, g in
               ((Tl, gl), (Tr, gr))]
        h = np.zeros(9)
        orc.thermo("species_enthalpy", 0.5 * (Tl + Tr), h)
        J = Fd[0][:, 0, 0, i]
        want = -0.5 * (lam[0] + lam[1]) * (Tr - Tl) / dx - (h * J).sum()
        got = F1[0][4, 0, 0, i] - F0[0][4, 0, 0, i]
        assert abs(got - want) <= 1e-9 * abs((h * J).sum()), (i, got, want)
        assert abs((h * J).sum()) > 0
        checked += 1
    assert checked == n + 1


def test_ebchem_diffusion_matches_the_regular_path(orc):
    """The EB-CHEM Fickian flux (FlowAdvanceBox.cpp, diffusion_{x,y,z}) is the
    regular path's (pinned by Fick's law above): in the cells of a cut box
    whose neighbours are all regular, the viscous minus the inviscid species
    dSdt of the two paths agree."""
    n, L = 16, 0.01
    rng = np.random.default_rng(9)
    boxes, A, As, dx, Y = _linear_composition_box(orc, n, 3.0, 800.0, 101325.0, L=L, dTdx=500.0)
    g = ocfd.make_geom((n, n, n), (0, 0, 0), (L, L, L), periodic=(1, 1, 1))
    sp = ocfd.eb_cylinder(0.5 * L, 0.5 * L, 0.22 * L)
    lo, hi = boxes[0]
    glo, ghi = orc.grown(lo, hi)
    gm = ocfd.eb_geometry(sp, g, lo, hi, 4)
    eb = orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"])
    par = orc.parm()
    dxi = [1 / dx] * 3
    _, e1 = orc.compute_dsdt_multi_box_eb(lo, hi, A[0], As[0], par, dxi, eb, 1)
    _, e0 = orc.compute_dsdt_multi_box_eb(lo, hi, A[0], As[0], par, dxi, eb, 0)
    _, r1 = orc.compute_dsdt_multi_box(lo, hi, A[0], As[0], par, dxi, 1)
    _, r0 = orc.compute_dsdt_multi_box(lo, hi, A[0], As[0], par, dxi, 0)
    reg = (eb.flag & 3) == 0
    ok = reg.copy()
    for ax in range(3):
        for sh in (-1, 1):
            ok &= np.roll(reg, sh, axis=ax)
    ok = ok[4:-4, 4:-4, 4:-4]
    assert ok.sum() > 100 and (~ok).sum() > 0
    de, dr = (e1 - e0)[:, ok], (r1 - r0)[:, ok]
    scale = np.abs(dr[[0, 8]]).max()
    assert scale > 0
    assert np.abs(de - dr).max() <= 1e-9 * scale
