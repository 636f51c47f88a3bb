"""Pins of the oracle's Rosenbrock (RODAS3) reactor integrator, which has no
counterpart in the reference (its chemistry takes fixed linearised-implicit
steps), by what the mathematics fixes:

* the complete fixed-temperature Jacobian of the production rates (mass
  action, third bodies, Troe fall-off) equals central differences of the
  rates, also where a species is absent;
* fixed steps converge to a tight scipy Radau solution of the same ODE as the
  step shrinks (a wrong tableau entry, sign or pivot exchange stalls this);
* at rtol 1e-4 the adaptive integration stays within its tolerance scale of
  that Radau solution and of the step-doubling integrator's 1 ms states, in at
  most 500 accepted steps per reactor on average.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp


def _reactor(orc, T0, p0=101325.0, phi=1.0):
    u, r = orc.reactor_init([T0], [p0], [phi])
    return u[0], r[0]


@pytest.mark.parametrize("T", [850.0, 1400.0, 2600.0])
def test_fixed_T_jacobian_matches_central_differences(orc, T):
    rng = np.random.default_rng(int(T))
    for trial in range(3):
        c = rng.uniform(1e-10, 2e-5, 9)
        c[3 + trial] = 0.0                       # an absent radical: exact derivative there too
        J = orc.jacobian_fixed_T(c, T)
        Jfd = np.zeros((9, 9))
        for j in range(9):
            d = max(1e-6 * c[j], 1e-13)
            cp, cm = c.copy(), c.copy()
            cp[j] += d
            cm[j] = c[j] - d if c[j] > d else c[j]
            wp, _ = orc.production_rate(cp, T)
            wm, _ = orc.production_rate(cm, T)
            Jfd[:, j] = (np.asarray(wp) - np.asarray(wm)) / (cp[j] - cm[j])
        err = np.abs(J - Jfd).max() / np.abs(Jfd).max()
        assert err < 1e-5, (T, trial, err)
        # the reference's mass-action matrix misses the third bodies and the
        # absent species' columns: the complete one must differ from it
        _, A = orc.production_rate(c, T)
        assert np.abs(np.asarray(A).reshape(9, 9) - Jfd).max() / np.abs(Jfd).max() > 10 * err


def _radau(orc, E0, y0, t_end, rtol=1e-11):
    f = lambda t, y: orc.reactor_rhs(E0, y0, y)[0]
    return solve_ivp(f, (0.0, t_end), y0, method="Radau", rtol=rtol, atol=1e-22).y[:, -1]


def test_fixed_steps_converge_to_radau(orc):
    u, r = _reactor(orc, 1200.0)
    ua, ra, _, _ = orc.reactor_adaptive(u[None], r[None], 5e-5, h0=1e-8, rtol=1e-8, atol=1e-14)
    E0, y0 = ua[0][4], ra[0].copy()           # in the ignition, with a radical pool
    span = 2e-6
    ref = _radau(orc, E0, y0, span)
    errs = []
    for N in (2, 4, 8, 16):
        uu = ua[0].copy()
        _, rr = orc.rodas3_fixed(uu, y0, span / N, N)
        errs.append(np.abs(rr - ref).max() / np.abs(ref).max())
    print("fixed-step errors", errs)
    for a, b in zip(errs, errs[1:]):
        assert b < a / 3.5                         # order above 1.8 on this stiff system
    assert errs[-1] < 2e-7


def test_adaptive_rosenbrock_within_tolerance_of_radau(orc):
    for T0, p0, phi in ((1000.0, 101325.0, 1.0), (1300.0, 5 * 101325.0, 0.6), (1500.0, 2 * 101325.0, 1.8)):
        u, r = _reactor(orc, T0, p0, phi)
        ub, rb, sb, jb = orc.reactor_rosenbrock(u[None], r[None], 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
        ref = _radau(orc, u[4], r, 1e-3, rtol=1e-9)
        _, Tref = orc.reactor_rhs(u[4], r, ref)
        Tb = orc.reactor_T(ub, rb)[0]
        print(T0, "steps", sb[0], "rejects", jb[0], "T", Tb, "Radau T", Tref)
        assert sb[0] > 0 and sb[0] <= 500
        assert abs(Tb - Tref) / Tref < 2e-4
        # the major species within the tolerance scale
        big = ref > 1e-3 * ref.max()
        assert np.abs(rb[0][big] - ref[big]).max() / np.abs(ref[big]).max() < 5e-4


def test_rosenbrock_agrees_with_step_doubling_in_fewer_steps(orc):
    import sys
    rng = np.random.default_rng(7)
    n = 48
    T0 = rng.uniform(900.0, 1500.0, n)
    p0 = rng.uniform(1.0, 10.0, n) * 101325.0
    phi = rng.uniform(0.5, 2.0, n)
    u, r = orc.reactor_init(T0, p0, phi)
    ua, ra, sa, _ = orc.reactor_adaptive(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    ub, rb, sb, jb = orc.reactor_rosenbrock(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    Ta, Tb = orc.reactor_T(ua, ra), orc.reactor_T(ub, rb)
    rel = np.abs(Ta - Tb) / Ta
    print("step doubling mean steps", sa.mean(), "RODAS3 mean steps", sb.mean(), "rejects", jb.mean(),
          "T rel diff max", rel.max(), file=sys.stderr)
    assert np.all(sb > 0) and sb.mean() <= 500 and sb.mean() < sa.mean() / 5
    assert rel.max() < 1e-3 and np.median(rel) < 1e-4
    # mass conserved to rounding (the production rates conserve it; clipping
    # at zero adds at most the clipped amounts)
    assert np.abs(rb.sum(axis=1) - ub[:, 0]).max() / ub[:, 0].max() < 1e-9


def test_step_with_a_nonphysical_stage_is_rejected(orc):
    """A lean, low-pressure reactor (1060.56 K, 1.074 atm, phi 0.530) whose
    induction-phase trial step sends a stage outside the thermodynamic tables
    (an extrapolated negative temperature, non-finite rates): the step is
    rejected and retried smaller, and the integration ends finite and close
    to Radau's."""
    T0, p0, phi = 1060.5603219620868, 108796.87196404164, 0.5295738294307426
    u, r = orc.reactor_init([T0], [p0], [phi])
    ub, rb, sb, jb = orc.reactor_rosenbrock(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    assert sb[0] > 0 and jb[0] >= 1
    assert np.isfinite(rb).all() and np.isfinite(ub).all()
    ref = _radau(orc, u[0][4], r[0], 1e-3, rtol=1e-9)
    _, Tref = orc.reactor_rhs(u[0][4], r[0], ref)
    Tb = orc.reactor_T(ub, rb)[0]
    assert abs(Tb - Tref) / Tref < 1e-3
