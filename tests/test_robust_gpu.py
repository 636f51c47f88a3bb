"""Robustness of the advance: the FP32-flux form on dimensional data with
strong jumps (the WENO-Z weights' squared ratios exceed single precision
there), the non-finite status of the device time control and of estTimeStep,
and the rule that a level refuses tile-shape changes while a captured step
graph holds its work-lists."""
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu


def _jump_state(lev, b, lo, hi, xjump, left, right, gamma=1.4):
    """Fill the whole grown FAB of box b with a z/y-uniform two-state x jump
    at cell xjump (dimensional: rho [kg/m^3], u [m/s], p [Pa])."""
    import torch
    nx = hi[0] - lo[0] + 1 + 8
    i = np.arange(nx) + lo[0] - 4
    out = np.empty((5,) + tuple(lev.fab_shape(b)[1:]))
    for n in range(5):
        out[n] = 0.0
    for side, (rho, u, p) in ((i < xjump, left), (i >= xjump, right)):
        out[0][..., side] = rho
        out[1][..., side] = rho * u
        out[4][..., side] = p / (gamma - 1.0) + 0.5 * rho * u * u
    return torch.from_numpy(out)


def _dimensional_level(nx=64, xjump=30, left=(10.0, 0.0, 1.0e6), right=(1.0, 0.0, 1.0e5)):
    import ocfd
    from ocfd import cases
    boxes, g, _ = cases.sod(nx, 8, 8)
    lev = ocfd.Level(boxes, g)
    S = lev.new_mf()
    for b, (lo, hi) in enumerate(boxes):
        lev.fab_view(S, b).copy_(_jump_state(lev, b, lo, hi, xjump, left, right))
    return boxes, g, lev, S


def _plain_fp32_weights_overflow(left, right, gamma=1.4):
    """Whether the plain single-precision WENO-Z weights of the energy f+ of
    this jump overflow: r = tau5/(1e-6 + beta_k) squared, with the split flux
    of the two states (Steger-Warming, u = 0: f+ energy = rho/(2 gamma) *
    (u+c) * ((u+c)^2/2 + (3-gamma)/(2(gamma-1)) c^2))."""
    def fpE(rho, u, p):
        c = np.sqrt(gamma * p / rho)
        lam = u + c
        return rho / (2 * gamma) * lam * (0.5 * lam * lam + (3 - gamma) / (2 * (gamma - 1)) * c * c)
    a, b = np.float32(fpE(*left)), np.float32(fpE(*right))
    V = np.array([a, a, a, b, b], dtype=np.float32)   # stencil of face f: cells f-3..f+1, jump before f
    t1, t2 = V[0] - 2 * V[1] + V[2], V[0] - 4 * V[1] + 3 * V[2]
    s11 = np.float32(13 / 12) * t1 * t1 + np.float32(0.25) * t2 * t2
    t5, t6 = V[2] - 2 * V[3] + V[4], 3 * V[2] - 4 * V[3] + V[4]
    s33 = np.float32(13 / 12) * t5 * t5 + np.float32(0.25) * t6 * t6
    r1 = np.abs(s11 - s33) / (np.float32(1e-6) + s11)
    with np.errstate(over="ignore"):
        return not np.isfinite(r1 * r1)


def test_fp32_flux_dimensional_jump_is_finite(torch_cuda, orc):
    """A dimensional pressure jump (10 kg/m^3, 1 MPa | 1 kg/m^3, 0.1 MPa):
    the FP32-flux dSdt is finite and within 1e-5 (relative to each
    component's largest magnitude) of the FP64 oracle's."""
    import ocfd
    left, right = (10.0, 0.0, 1.0e6), (1.0, 0.0, 1.0e5)
    assert _plain_fp32_weights_overflow(left, right)   # the case the scaled weights exist for
    boxes, g, lev, S = _dimensional_level(left=left, right=right)
    p = ocfd.default_parm()
    out = {}
    for name, flags in (("fp64", 0), ("fp32", ocfd.FLUX_FP32)):
        D = lev.new_mf(ngrow=0)
        lev.compute_dsdt(S, D, p, flags)
        out[name] = [lev.fab_view(D, b, 0).cpu().numpy() for b in range(lev.nbox)]
    # FP64 against the oracle
    dx = g.dx
    dxinv = [1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]]
    fabs = orc.new_level_fabs(boxes)
    for b, (lo, hi) in enumerate(boxes):
        fabs[b].a[...] = lev.fab_view(S, b).cpu().numpy()
        ref, _ = orc.compute_dsdt_box(lo, hi, fabs[b], orc.parm(), dxinv)
        got64, got32 = out["fp64"][b], out["fp32"][b]
        assert np.isfinite(got32).all()
        scale = np.abs(ref).reshape(5, -1).max(axis=1)
        scale[1:4] = scale[1:4].max()          # momentum components share the vector's scale
        scale = scale.reshape(5, 1, 1, 1)
        assert (np.abs(got64 - ref) <= 1e-10 * scale).all()
        err32 = (np.abs(got32 - ref) / scale).max()
        print("fp32 vs fp64 oracle, max rel to component max:", err32)
        assert err32 < 1e-5


def test_fp32_flux_dimensional_shock_tube_steps(torch_cuda):
    """Twenty device-driven SSP-RK3 steps of the dimensional shock tube in the
    FP32-flux form: every state finite, status OK, and the density within
    1e-5 relative L1 of the FP64 form."""
    import ocfd
    p = ocfd.default_parm()
    res = {}
    for name, flags in (("fp64", 0), ("fp32", ocfd.FLUX_FP32)):
        boxes, g, lev, A = _dimensional_level()
        B, C = lev.new_mf(), lev.new_mf()
        ctl = ocfd.new_ctl()
        lev.ctl_start(A, ctl, p, 0.5, max_steps=20)
        for _ in range(20):
            lev.step_device(A, B, C, ctl, p, 2, flags)
        c = ocfd.read_ctl(ctl)
        assert c.status == 0 and c.steps == 20
        res[name] = [a[:, 4:-4, 4:-4, 4:-4] for a in gu.download(lev, A)]
    for a in res["fp32"]:
        assert np.isfinite(a).all()
    e = gu.level_rel_err(res["fp32"], res["fp64"])
    print("fp32 vs fp64 after 20 steps (L1, Linf) per component", e)
    assert e[0][0] < 1e-5


def test_nonfinite_state_stops_device_run(torch_cuda):
    """A NaN in one valid cell: estTimeStep reports OCFD_ENONFINITE, and the
    device time control ends the run with that status instead of stepping
    on with the NaN dropped from the CFL minimum."""
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.sod(64, 8, 8)
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    lev.init_state(kind, A, p)
    lev.fill_boundary(A, 0.0)
    lev.estdt(A, 0.5, p)   # finite state: no error
    lev.fab_view(A, 0)[4, 6, 6, 20] = float("nan")   # energy of a valid cell
    with pytest.raises(RuntimeError, match="status 5"):
        lev.estdt(A, 0.5, p)
    ctl = ocfd.new_ctl()
    lev.ctl_start(A, ctl, p, 0.5, max_steps=10)
    for _ in range(3):
        lev.step_device(A, B, C, ctl, p, 2, 0)
    c = ocfd.read_ctl(ctl)
    assert c.done == 1 and c.status == ocfd.ENONFINITE and c.steps == 0


def test_nonfinite_after_a_step_sets_status(torch_cuda):
    """The reduction folded into the finishing sweep flags a non-finite new
    state: a NaN planted in U^n's valid cell propagates into the step's
    result, and the dt update then ends the run with OCFD_ENONFINITE."""
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.kh3d(32)
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    lev.init_state(kind, A, p, seed=3)
    ctl = ocfd.new_ctl()
    lev.ctl_start(A, ctl, p, 0.5, max_steps=10)
    lev.step_device(A, B, C, ctl, p, 2, ocfd.VISC | ocfd.FLUX_FP32)
    assert ocfd.read_ctl(ctl).status == 0
    lev.fab_view(A, 0)[0, 10, 10, 10] = float("nan")
    lev.step_device(A, B, C, ctl, p, 2, ocfd.VISC | ocfd.FLUX_FP32)
    c = ocfd.read_ctl(ctl)
    assert c.done == 1 and c.status == ocfd.ENONFINITE and c.steps == 2


def test_graph_holds_the_tile_shape(torch_cuda):
    """While a captured step graph exists the level refuses a new march chunk,
    the autotuner and its own destruction; once the graph is gone the chunk
    changes, and a graph captured then replays bit-equal to eager steps."""
    import torch
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.kh3d(32)
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    flags = ocfd.VISC | ocfd.FLUX_FP32
    s = torch.cuda.Stream()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    lev.init_state(kind, A, p, seed=5)
    A0 = A.clone()
    ctl = ocfd.new_ctl()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        lev.ctl_start(A, ctl, p, 0.5, max_steps=4, stream=s)
        G = lev.capture_steps(A, B, C, ctl, p, 2, 2, flags, stream=s)
    ch = lev.march_chunk()
    with pytest.raises(RuntimeError, match="graphs"):
        lev.march_chunk(8 if ch != 8 else 16)
    lev.fill_boundary(A0, 0.0)
    with pytest.raises(RuntimeError, match="graphs"):
        lev.autotune(A0, p, flags, reps=1)
    assert ocfd.lib().ocfd_level_destroy(lev.handle) == ocfd.EINVAL
    assert lev.march_chunk() == ch
    del G
    assert lev.march_chunk(8 if ch != 8 else 16) != ch
    # eager reference and a graph captured after the change
    E, Eb, Ec = A0.clone(), lev.new_mf(), lev.new_mf()
    ectl = ocfd.new_ctl()
    lev.ctl_start(E, ectl, p, 0.5, max_steps=4)
    for _ in range(4):
        lev.step_device(E, Eb, Ec, ectl, p, 2, flags)
    A.copy_(A0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        lev.ctl_start(A, ctl, p, 0.5, max_steps=4, stream=s)
        G = lev.capture_steps(A, B, C, ctl, p, 2, 2, flags, stream=s)
        G.launch(s)
        G.launch(s)
    s.synchronize()
    for a, b in zip(gu.download(lev, A), gu.download(lev, E)):
        assert np.array_equal(a[:, 4:-4, 4:-4, 4:-4], b[:, 4:-4, 4:-4, 4:-4])
    del G


def test_device_pool_reuses_level_blocks(torch_cuda):
    """A destroyed level's tables and scratch go back to the library's pool and
    a level of the same boxes takes them again; its dSdt equals that of a
    level built on fresh blocks; releasing the pool empties it."""
    import gc
    import torch
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.kh3d(32, max_grid=(16, 16, 32))
    p = ocfd.default_parm()
    ocfd.release_cached_memory()
    assert ocfd.cached_device_bytes() == 0
    lev = ocfd.Level(boxes, g)
    S = lev.new_mf()
    lev.init_state(kind, S, p, seed=11)
    D1 = lev.new_mf(ngrow=0)
    lev.compute_dsdt(S, D1, p, ocfd.VISC)
    torch.cuda.synchronize()
    del lev
    gc.collect()
    held = ocfd.cached_device_bytes()
    assert held > 0
    lev2 = ocfd.Level(boxes, g)
    assert ocfd.cached_device_bytes() < held          # blocks taken from the pool
    D2 = lev2.new_mf(ngrow=0)
    lev2.compute_dsdt(S, D2, p, ocfd.VISC)
    torch.cuda.synchronize()
    assert torch.equal(D1, D2)
    del lev2
    gc.collect()
    ocfd.release_cached_memory()
    assert ocfd.cached_device_bytes() == 0
