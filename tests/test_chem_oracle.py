"""Pins of the chemistry oracle (Li-Dryer H2/air) against the reference's own
AMReX-free kinetics (Test/chemical_test.cpp built into oracle/_ref), the
reference's known answers (Exec/7_JET/prob.cpp:18-37), conservation laws,
numpy's linear solver and a library stiff ODE solve.  (The reference's
Jacobian holds third-body efficiencies and fall-off factors fixed, so it is
not the full derivative of wdot; it is pinned bitwise by the reference build
and by element conservation of its columns.)"""
import ctypes

import numpy as np
import pytest

# atoms (H, O, N) per species, order H2 O2 H2O H O OH HO2 H2O2 N2
ATOMS = np.array([[2, 0, 0], [0, 2, 0], [2, 1, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [1, 2, 0], [2, 2, 0], [0, 0, 2]])


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _random_state(rng, n=1):
    """Random mixtures (mol/cm^3) and temperatures across both NASA7 branches."""
    c = rng.uniform(0, 1, (n, 9)) ** 3 * 1e-5
    c[:, 8] += 2e-5
    T = rng.uniform(600, 2800, n)
    return c, T


@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref_chem()
    if r is None:
        pytest.skip("oracle/_ref/libchemref.so not built (needs the reference sources)")
    return r


def test_production_rate_matches_reference_build(orc, ref):
    rng = np.random.default_rng(0)
    c, T = _random_state(rng, 200)
    for m in range(len(T)):
        w, J = orc.production_rate(c[m], T[m])
        wr = np.zeros(9); Jr = np.zeros(81)
        cm = np.ascontiguousarray(c[m])
        ref.vproductionRate(_dp(wr), _dp(Jr), _dp(cm), T[m])
        assert np.array_equal(w, wr), m
        assert np.array_equal(J.ravel(), Jr), m


def test_thermo_matches_reference_build(orc, ref):
    rng = np.random.default_rng(1)
    for _ in range(100):
        rhoi = rng.uniform(0, 1, 9) ** 2
        T = rng.uniform(300, 3400)
        out = ctypes.c_double()
        for nm in ("CKUBMS", "CKCVBS", "CKCPBS", "CKGAMMA"):
            getattr(ref, nm)(T, _dp(rhoi), ctypes.byref(out))
            assert orc.thermo(nm, T, rhoi) == out.value, nm
        ref.CKPY(_dp(rhoi), T, ctypes.byref(out))
        assert orc.thermo("CKPY", rhoi, T) == out.value
        e = orc.thermo("CKUBMS", T, rhoi) * rng.uniform(0.9, 1.1)
        Tg, _ = orc.T_given_ey(e, rhoi)
        Tr = ctypes.c_double()
        ref.GET_T_GIVEN_EY(e, _dp(rhoi.copy()), ctypes.byref(Tr))
        assert Tg == Tr.value
        y = rhoi / rhoi.sum()
        x = np.zeros(9)
        ref.CKYTX(_dp(y), _dp(x))
        assert np.array_equal(orc.CKYTX(y), x)
        ref.CKXTY(_dp(x), _dp(y))
        assert np.array_equal(orc.CKXTY(x), y)


def test_gauss_matches_reference_build_and_numpy(orc, ref):
    rng = np.random.default_rng(2)
    for _ in range(100):
        A = np.eye(9) + rng.normal(0, 1, (9, 9)) * rng.uniform(0.01, 3)
        b = rng.normal(0, 1, 9)
        x = orc.gauss(A, b)
        xr = np.zeros(9)
        ref.gauss(_dp(xr), _dp(np.ascontiguousarray(A)), _dp(b))
        assert np.array_equal(x, xr)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


def test_jet_known_answers(orc):
    """Exec/7_JET/prob.cpp:18-37: jet (X_H2 .15, X_N2 .85, 305 K, 1 atm) and
    ambient (X_O2 .21, X_N2 .79, 1150 K, 1 atm) densities and sensible energies."""
    for X, T, rho_ref, e_ref, y in (
            ({0: 0.15, 8: 0.85}, 305.0, 0.963512265437802995, 245428.076382951345, (0, 0.012540)),
            ({1: 0.21, 8: 0.79}, 1150.0, 0.305731395505996983, 271976.980564150028, (1, 0.232909))):
        x = np.zeros(9)
        for k, v in X.items():
            x[k] = v
        Y = orc.CKXTY(x)
        rho = orc.thermo("CKRHOY", 101325.0, T, Y)
        assert abs(rho - rho_ref) < 1e-14 * rho_ref
        e = orc.thermo("CKUBMS", T, Y * rho)
        assert abs(e - e_ref) < 1e-9 * e_ref
        assert abs(Y[y[0]] - y[1]) < 5e-7


def test_mass_and_element_conservation(orc):
    mw, _, _ = orc.chem_tables()
    rng = np.random.default_rng(3)
    c, T = _random_state(rng, 300)
    for m in range(len(T)):
        w, J = orc.production_rate(c[m], T[m])
        scale = np.abs(w * mw).max() + 1e-300
        assert abs((w * mw).sum()) < 1e-12 * scale
        assert np.all(np.abs(ATOMS.T @ w) < 1e-12 * (np.abs(w).max() + 1e-300))
        # the Jacobian columns conserve atoms too
        assert np.all(np.abs(ATOMS.T @ J) < 1e-10 * (np.abs(J).max() + 1e-300))


def test_T_inversion_round_trip(orc):
    rng = np.random.default_rng(5)
    for _ in range(100):
        rhoi = rng.uniform(0, 1, 9) ** 2
        T = rng.uniform(260, 3400)
        e = orc.thermo("CKUBMS", T, rhoi)
        Tg, it = orc.T_given_ey(e, rhoi)
        assert abs(Tg - T) < 1e-6 and it < 200


def test_chem_step_conserves_mass_and_total_energy(orc):
    """rho_k sums stay at rho and the sensible+formation energy is conserved."""
    mw, lo, _ = orc.chem_tables()
    T0 = np.array([1200.0, 1000.0]); p0 = np.array([101325.0, 5 * 101325.0]); phi = np.array([1.0, 0.5])
    u, r = orc.reactor_init(T0, p0, phi)
    Ru = 8.31446261815324
    hf = lo[:, 5] * Ru / mw   # formation part of the energy, J/kg
    tot0 = u[:, 4] + r @ hf
    for _ in range(20):
        u, r = orc.chem_step(u, r, 1e-6)
    assert np.allclose(r.sum(1), u[:, 0], rtol=1e-12)
    assert np.allclose(u[:, 4] + r @ hf, tot0, rtol=1e-10)


def test_adaptive_integration_matches_library_stiff_solver(orc):
    """Constant-volume ignition to 1 ms: the step-doubling integrator agrees
    with scipy's Radau on the same right-hand side (the reference's rates,
    d rho_k/dt = W_k wdot_k 1e6, de/dt = -sum_k h_f,k d rho_k/dt)."""
    from scipy.integrate import solve_ivp
    mw, lo, _ = orc.chem_tables()
    Ru = 8.31446261815324
    hf = lo[:, 5] * Ru / mw
    u0, r0 = orc.reactor_init([1100.0], [101325.0], [1.0])

    def rhs(t, y):
        rhoi = np.maximum(y[:9], 0.0)
        T, _ = orc.T_given_ey(y[9], rhoi)
        w, _ = orc.production_rate(rhoi / mw * 1e-6, T)
        dr = w * mw * 1e6
        return np.concatenate([dr, [-(dr @ hf)]])

    def jac(t, y):
        rhoi = np.maximum(y[:9], 0.0)
        T, _ = orc.T_given_ey(y[9], rhoi)
        _, J = orc.production_rate(rhoi / mw * 1e-6, T)
        Jm = J * mw[:, None] / mw[None, :]
        out = np.zeros((10, 10))
        out[:9, :9] = Jm
        out[9, :9] = -(hf @ Jm)
        return out

    y0 = np.concatenate([r0[0], [u0[0, 4]]])
    sol = solve_ivp(rhs, (0, 1e-3), y0, method="Radau", jac=jac, rtol=1e-10, atol=1e-14)
    assert sol.success
    u, r, steps, rej = orc.reactor_adaptive(u0, r0, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    assert steps[0] > 0
    T_lib, _ = orc.T_given_ey(sol.y[9, -1], np.maximum(sol.y[:9, -1], 0))
    T_ad = orc.reactor_T(u, r)[0]
    assert T_lib > 2000.0                              # ignited
    assert abs(T_ad - T_lib) < 5e-5 * T_lib, (T_ad, T_lib)
    yw = r[0] / r[0].sum()
    yl = sol.y[:9, -1] / sol.y[:9, -1].sum()
    assert abs(yw[2] - yl[2]) < 2e-4 * yl[2]           # H2O


def test_renorm(orc):
    rng = np.random.default_rng(6)
    u = np.zeros((50, 5)); u[:, 0] = rng.uniform(0.1, 2, 50)
    r = rng.uniform(-0.01, 1, (50, 9))
    out = orc.renorm(u, r)
    assert np.all(out >= 0)
    assert np.allclose(out.sum(1), u[:, 0], rtol=1e-14)
    clipped = np.maximum(r, 0)
    assert np.allclose(out, clipped * (u[:, 0] / clipped.sum(1))[:, None], rtol=1e-14)


@pytest.fixture(scope="module")
def reft(orc):
    r = orc.ref_trans()
    if r is None:
        pytest.skip("oracle/_ref/libtransref.so not built (needs the reference sources)")
    return r


def test_transport_matches_reference_build(orc, reft):
    """mixtureProperties / getMixDiffCoeffsMass bitwise against the reference's
    own AMReX-free copy (Test/transport_test.cpp)."""
    rng = np.random.default_rng(8)
    for _ in range(200):
        X = rng.uniform(0, 1, 9) ** 2
        X /= X.sum()
        T = rng.uniform(250, 3000)
        p = rng.uniform(0.1, 20) * 101325.0
        mu, lam = orc.mixture_properties(T, X)
        mr = ctypes.c_double(); lr = ctypes.c_double()
        Xc = np.ascontiguousarray(X)
        reft.mixtureProperties(T, _dp(Xc), ctypes.byref(mr), ctypes.byref(lr))
        assert mu == mr.value and lam == lr.value
        D = orc.mix_diff_coeffs(T, p, X)
        Dr = np.zeros(9)
        reft.getMixDiffCoeffsMass(T, p, _dp(Xc), _dp(Dr))
        assert np.array_equal(D, Dr)


def test_transport_limits(orc):
    """A pure species gets its own viscosity and conductivity (Wilke's
    phi_kk = 1, lambda = (X lambda + 1/(X/lambda))/2); air-like mixtures at
    300 K have mu ~ 1.8e-5 Pa s and lambda ~ 0.026 W/m/K; D scales as 1/p."""
    T = 300.0
    mus = []
    for k in range(9):
        X = np.zeros(9); X[k] = 1.0
        mu, lam = orc.mixture_properties(T, X)
        assert mu > 0 and lam > 0
        mus.append(mu)
    X = np.zeros(9); X[1] = 0.21; X[8] = 0.79
    mu, lam = orc.mixture_properties(T, X)
    assert 1.6e-5 < mu < 2.0e-5 and 0.022 < lam < 0.030
    X2 = X.copy(); X2[0] = 1e-3; X2 /= X2.sum()
    D1 = orc.mix_diff_coeffs(T, 101325.0, X2)
    D2 = orc.mix_diff_coeffs(T, 2 * 101325.0, X2)
    assert np.allclose(D1, 2 * D2, rtol=1e-14)
    assert 5e-5 < D1[0] < 1.5e-4          # H2 in air ~ 7.8e-5 m^2/s at 300 K
