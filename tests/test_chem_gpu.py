"""GPU parity of the chemistry kernels against the chemistry oracle.

Fixed-step paths (the reference's linearised-implicit step) are compared per
reactor/cell at the relative tolerances written below; both sides use FP64
in the reference's operation order, so differences come only from the device
libm (exp/log/pow) and can move the Newton temperature exit by one iteration.
The adaptive integrator is compared at its own tolerance scale.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def _oracle_state(orc, T0, p0, phi):
    u, r = orc.reactor_init(T0, p0, phi)
    return u, r


def _gpu_rows(R):
    a = R.rows().cpu().numpy()
    return a[0], a[1], a[2:].T.copy()


def test_reactor_init_matches_oracle(torch_cuda, orc):
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(512, seed=1)
    R = ocfd.Reactors(T0, p0, phi)
    rho, E, Y = _gpu_rows(R)
    u, r = _oracle_state(orc, T0, p0, phi)
    assert np.allclose(rho, u[:, 0], rtol=1e-15, atol=0)
    assert np.allclose(E, u[:, 4], rtol=1e-14, atol=0)
    assert np.allclose(Y, r, rtol=1e-15, atol=0)
    T = R.temperature().cpu().numpy()
    assert np.allclose(T, T0, rtol=1e-9)


@pytest.mark.parametrize("dt,nsteps", [(1e-5, 100), (1e-6, 100), (1e-7, 200)])
def test_reactor_fixed_steps_match_oracle(torch_cuda, orc, dt, nsteps):
    """4096 seeded reactors, nsteps reference steps of dt.  The reference's
    fixed step is ill-conditioned for part of the batch at large dt (a 1e-14
    perturbation of the initial energy changes the final state, or it
    diverges): those reactors are identified with the oracle itself and
    excluded; the rest (>= 90 %) must agree with the oracle to 1e-8."""
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(4096, seed=7)
    R = ocfd.Reactors(T0, p0, phi)
    R.advance_fixed(dt, nsteps)
    rho, E, Y = _gpu_rows(R)
    u0, r0 = _oracle_state(orc, T0, p0, phi)
    u, r = orc.reactor_fixed(u0, r0, dt, nsteps)
    up = u0.copy()
    up[:, 4] *= 1.0 + 1e-14
    ub, rb = orc.reactor_fixed(up, r0, dt, nsteps)
    To = orc.reactor_T(u, r)
    Tb = orc.reactor_T(ub, rb)
    Tg = R.temperature().cpu().numpy()
    with np.errstate(invalid="ignore"):
        well = np.isfinite(To) & np.isfinite(Tb) & (np.abs(To - Tb) < 1e-8 * np.abs(To))
    assert well.mean() >= 0.9
    Yw, rw, Ew, uEw = Y[well], r[well], E[well], u[well, 4]
    err = np.abs(Yw - rw).sum(0) / np.abs(rw).sum(0).clip(1e-300)
    errE = np.abs(Ew - uEw).sum() / np.abs(uEw).sum()
    relT = np.abs(Tg[well] - To[well]) / To[well]
    print(f"dt {dt}: well-conditioned {well.mean():.4f}; rel L1 per species", err, "E", errE, "T max rel",
          relT.max())
    assert err.max() < 1e-8 and errE < 1e-10 and relT.max() < 1e-8


def test_reactor_adaptive_matches_oracle(torch_cuda, orc):
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(512, seed=9)
    R = ocfd.Reactors(T0, p0, phi)
    R.advance_adaptive(1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    Tg = R.temperature().cpu().numpy()
    steps = R.steps.cpu().numpy()
    u, r = _oracle_state(orc, T0, p0, phi)
    u, r, so, ro = orc.reactor_adaptive(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    To = orc.reactor_T(u, r)
    assert np.all(steps > 0) and np.all(so > 0)
    rel = np.abs(Tg - To) / To
    print("adaptive T rel diff: median", np.median(rel), "max", rel.max(), "same step counts", np.mean(steps == so))
    assert np.median(rel) < 1e-8 and rel.max() < 1e-4


def test_level_chem_advance_and_renorm(torch_cuda, orc):
    import torch
    import ocfd
    from ocfd import cases
    boxes, g, _ = cases.sod(32, 8, 8, max_grid=(16, 8, 8))
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    rng = np.random.default_rng(3)
    n = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in boxes)
    T0 = rng.uniform(1000, 1400, n); p0 = rng.uniform(1, 5, n) * 101325.0; phi = rng.uniform(0.5, 2, n)
    u, r = _oracle_state(orc, T0, p0, phi)
    vel = rng.uniform(-50, 50, (n, 3))
    u[:, 1:4] = u[:, :1] * vel
    u[:, 4] += 0.5 * u[:, 0] * (vel ** 2).sum(1)
    S = lev.new_mf(4, 5)
    Sp = lev.new_mf(4, 9)
    # scatter the cells into the level MultiFabs (valid cells, box order, i fastest)
    m = 0
    for b, (lo, hi) in enumerate(boxes):
        nb = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
        shp = (hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        fs = lev.fab_view(S, b, 4, 5)
        fp = lev.fab_view(Sp, b, 4, 9)
        fs[:, 4:-4, 4:-4, 4:-4] = torch.from_numpy(u[m:m + nb].T.reshape((5,) + shp).copy())
        fp[:, 4:-4, 4:-4, 4:-4] = torch.from_numpy(r[m:m + nb].T.reshape((9,) + shp).copy())
        m += nb
    lev.chem_advance(S, Sp, 2e-6, p)
    uo, ro = orc.chem_step(u, r, 2e-6)
    got_u = []; got_r = []
    for b, (lo, hi) in enumerate(boxes):
        got_u.append(lev.fab_view(S, b, 4, 5)[:, 4:-4, 4:-4, 4:-4].reshape(5, -1).cpu().numpy().T)
        got_r.append(lev.fab_view(Sp, b, 4, 9)[:, 4:-4, 4:-4, 4:-4].reshape(9, -1).cpu().numpy().T)
    gu_ = np.concatenate(got_u); gr = np.concatenate(got_r)
    assert np.max(np.abs(gr - ro) / (np.abs(ro).max(0) + 1e-300)) < 1e-10
    assert np.max(np.abs(gu_ - uo) / (np.abs(uo).max(0) + 1e-300)) < 1e-12
    # renormalisation after perturbing the species
    for b in range(len(boxes)):
        v = lev.fab_view(Sp, b, 4, 9)
        v[:, 4:-4, 4:-4, 4:-4] *= 1.0 + 0.01 * torch.rand_like(v[:, 4:-4, 4:-4, 4:-4])
    before = np.concatenate([lev.fab_view(Sp, b, 4, 9)[:, 4:-4, 4:-4, 4:-4].reshape(9, -1).cpu().numpy().T
                             for b in range(len(boxes))])
    lev.species_renorm(S, Sp)
    after = np.concatenate([lev.fab_view(Sp, b, 4, 9)[:, 4:-4, 4:-4, 4:-4].reshape(9, -1).cpu().numpy().T
                            for b in range(len(boxes))])
    ref = orc.renorm(gu_, before)
    assert np.array_equal(after, ref)


def test_reactor_adaptive_lane_modes_agree(torch_cuda):
    """Small batches run two lanes per reactor, full batches one: the same
    reactors give the same states, step and rejection counts bit for bit."""
    import numpy as np
    import torch
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(256, seed=13)
    small = ocfd.Reactors(T0, p0, phi)
    small.advance_adaptive(2e-4, h0=1e-8, rtol=1e-4, atol=1e-9)
    n = 64 * 4 * torch.cuda.get_device_properties(0).multi_processor_count   # one wave per SIMD: one lane each
    reps = -(-n // 256)
    big = ocfd.Reactors(np.tile(T0, reps)[:n], np.tile(p0, reps)[:n], np.tile(phi, reps)[:n])
    big.advance_adaptive(2e-4, h0=1e-8, rtol=1e-4, atol=1e-9)
    a = small.rows().cpu().numpy()
    b = big.rows().cpu().numpy()[:, :256]
    assert np.array_equal(a, b)
    assert np.array_equal(small.steps.cpu().numpy(), big.steps.cpu().numpy()[:256])
    assert np.array_equal(small.rejects.cpu().numpy(), big.rejects.cpu().numpy()[:256])


def test_reactor_chunks_equal_one_launch(torch_cuda):
    """The resumable integration (a few substep attempts per reactor per
    launch) gives the one-launch integration's states, step and rejection
    counts bit for bit, and reports no reactor remaining at the end."""
    import numpy as np
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(300, seed=21)
    one = ocfd.Reactors(T0, p0, phi)
    one.advance_adaptive(2e-4, h0=1e-8, rtol=1e-4, atol=1e-9)
    ch = ocfd.Reactors(T0, p0, phi)
    ch.begin_chunks(1e-8)
    launches = 0
    while True:
        ch.chunk(2e-4, budget=7, rtol=1e-4, atol=1e-9)
        launches += 1
        if int(ch.remaining.item()) == 0:
            break
        assert launches < 100000
    assert launches > 1
    assert np.array_equal(one.rows().cpu().numpy(), ch.rows().cpu().numpy())
    assert np.array_equal(one.steps.cpu().numpy(), ch.steps.cpu().numpy())
    assert np.array_equal(one.rejects.cpu().numpy(), ch.rejects.cpu().numpy())


def test_reactor_rosenbrock_matches_oracle(torch_cuda, orc):
    """The RODAS3 reactor kernel against the oracle's identical integrator:
    temperatures at 1 ms within the tolerance scale (the device libm can move
    a step decision), most step counts equal, and at most 500 accepted steps
    per reactor on average."""
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(512, seed=9)
    R = ocfd.Reactors(T0, p0, phi)
    R.advance_rosenbrock(1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    Tg = R.temperature().cpu().numpy()
    steps = R.steps.cpu().numpy()
    u, r = _oracle_state(orc, T0, p0, phi)
    u, r, so, ro = orc.reactor_rosenbrock(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    To = orc.reactor_T(u, r)
    assert np.all(steps > 0) and np.all(so > 0)
    rel = np.abs(Tg - To) / To
    print("rosenbrock T rel diff: median", np.median(rel), "max", rel.max(), "same step counts",
          np.mean(steps == so), "mean steps", steps.mean())
    assert np.median(rel) < 1e-8 and rel.max() < 1e-4
    assert np.mean(steps == so) > 0.9
    assert steps.mean() <= 500
    # a reactor whose trial step leaves the thermodynamic tables: rejected
    # and retried, the integration ends finite
    one = ocfd.Reactors([1060.5603219620868], [108796.87196404164], [0.5295738294307426])
    one.advance_rosenbrock(1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    assert int(one.steps[0]) > 0 and int(one.rejects[0]) >= 1
    assert np.isfinite(one.rows().cpu().numpy()).all()


def test_rosenbrock_chunks_equal_one_launch(torch_cuda):
    """The resumable RODAS3 integration gives the one-launch states, step and
    rejection counts bit for bit."""
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(300, seed=23)
    one = ocfd.Reactors(T0, p0, phi)
    one.advance_rosenbrock(5e-4, h0=1e-8, rtol=1e-4, atol=1e-9)
    ch = ocfd.Reactors(T0, p0, phi)
    ch.begin_chunks(1e-8, method="rosenbrock")
    launches = 0
    while True:
        ch.chunk(5e-4, budget=5, rtol=1e-4, atol=1e-9)
        launches += 1
        if int(ch.remaining.item()) == 0:
            break
        assert launches < 100000
    assert launches > 1
    assert np.array_equal(one.rows().cpu().numpy(), ch.rows().cpu().numpy())
    assert np.array_equal(one.steps.cpu().numpy(), ch.steps.cpu().numpy())
    assert np.array_equal(one.rejects.cpu().numpy(), ch.rejects.cpu().numpy())
