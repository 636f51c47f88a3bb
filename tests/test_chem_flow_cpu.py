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
