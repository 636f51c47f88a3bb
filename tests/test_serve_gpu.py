"""Device-driven time steps, their hipGraph replay and the request loop.

The device-driven step applies the driver's dt rule (Source/EBR.cpp:189-306)
on the device; it must reproduce the host-driven loop bit for bit.  The
request loop must give every case the same result alone and in a mix of
concurrently resident cases."""
import pytest
import torch

import ocfd
from ocfd import cases, serve

pytestmark = pytest.mark.gpu


def _setup(case, seed=0):
    boxes, g, kind = case
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    A = lev.new_mf()
    lev.init_state(kind, A, p, seed)
    return lev, p, A


def _device_run(lev, p, A0, cfl, scheme, flags, stop_time, max_steps, nsteps):
    A = A0.clone(); B = torch.zeros_like(A); C = torch.zeros_like(A)
    ctl = ocfd.new_ctl()
    lev.ctl_start(A, ctl, p, cfl, 0.0, stop_time, max_steps)
    for _ in range(nsteps):
        lev.step_device(A, B, C, ctl, p, scheme, flags)
    torch.cuda.synchronize()
    return A, ocfd.read_ctl(ctl)


CASES = [
    ("sod", lambda: cases.sod(64, 8, 4), 0.5, 2, 0, 0.05, -1),
    ("sod_fp32_planar", lambda: cases.sod(64, 8, 2), 0.5, 2, ocfd.FLUX_FP32 | ocfd.PLANAR, 0.05, -1),
    ("dmr", lambda: cases.dmr(64, 16, 2, sizes=(8, 16), seed=3), 0.5, 2, ocfd.PLANAR, -1.0, 6),
    ("kh_rk2", lambda: cases.kh3d(16), 0.5, 1, ocfd.VISC, -1.0, 4),
    ("kh_euler_fp32", lambda: cases.kh3d(16), 0.3, 0, ocfd.VISC | ocfd.FLUX_FP32, -1.0, 3),
]


@pytest.mark.parametrize("name,make,cfl,scheme,flags,stop,maxs", CASES, ids=[c[0] for c in CASES])
def test_device_driven_equals_host_driven(torch_cuda, name, make, cfl, scheme, flags, stop, maxs):
    lev, p, A0 = _setup(make(), seed=5)
    A = A0.clone(); B = torch.zeros_like(A); C = torch.zeros_like(A)
    steps, t, dt = ocfd.run(lev, A, B, C, p, cfl, scheme=scheme, flags=flags, stop_time=stop, nsteps=maxs)
    torch.cuda.synchronize()
    Ad, c = _device_run(lev, p, A0, cfl, scheme, flags, stop, maxs, steps + 3)   # 3 extra steps are no-ops
    assert c.done == 1 and c.steps == steps
    assert c.time == t and c.dt == dt
    assert torch.equal(Ad, A), (Ad - A).abs().max().item()
    if stop > 0:
        assert t == stop


def test_graph_replay_equals_eager(torch_cuda):
    lev, p, A0 = _setup(cases.kh3d(16), seed=9)
    flags = ocfd.VISC | ocfd.FLUX_FP32
    Ae, ce = _device_run(lev, p, A0, 0.5, 2, flags, -1.0, 10, 10)
    A = A0.clone(); B = torch.zeros_like(A); C = torch.zeros_like(A)
    ctl = ocfd.new_ctl()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        lev.ctl_start(A, ctl, p, 0.5, 0.0, -1.0, 10, stream=s)
        g = lev.capture_steps(A, B, C, ctl, p, 4, 2, flags, stream=s)
    other = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
    for r in range(3):   # 12 steps: the last two are no-ops (max_steps 10)
        g.launch(s)
        s.synchronize()
        other.mul_(1.5).add_(1.0)   # eager work between replays
        torch.cuda.synchronize()
        c = ocfd.read_ctl(ctl)
        assert c.steps == min(4 * (r + 1), 10)
    assert c.done == 1 and c.time == ce.time and c.dt == ce.dt
    assert torch.equal(A, Ae)


def test_footprint_covers_device_memory(torch_cuda):
    for make in (lambda: cases.kh3d(32), lambda: cases.dmr(128, 32, 2, sizes=(16, 32), seed=4)):
        req = serve.FlowRequest("x", make)
        boxes, g, kind = make()
        lev = ocfd.Level(boxes, g)
        actual = lev.device_bytes() + 3 * 8 * lev.mf_size() + ocfd.CTL_BYTES
        fp = req.footprint()
        assert actual <= fp <= actual + (2 << 20)


def _mix():
    F = ocfd.FLUX_FP32
    return [
        serve.FlowRequest("sod64", lambda: cases.sod(64, 8, 2), flags=F | ocfd.PLANAR, stop_time=0.05),
        serve.FlowRequest("sod128", lambda: cases.sod(128, 16, 2, max_grid=(64, 64, 2)), flags=F | ocfd.PLANAR,
                          stop_time=0.02),
        serve.FlowRequest("dmr", lambda: cases.dmr(128, 32, 2, sizes=(16, 32), seed=11), flags=F | ocfd.PLANAR,
                          stop_time=0.002),
        serve.FlowRequest("kh32", lambda: cases.kh3d(32), flags=F | ocfd.VISC, max_steps=5, seed=42),
        serve.FlowRequest("kh16", lambda: cases.kh3d(16), flags=ocfd.VISC, max_steps=7, seed=1),
        serve.FlowRequest("chemeb", lambda: cases.chem_channel(64, 16, 2, max_grid=16), flags=ocfd.VISC, cfl=20.0,
                          scheme=1, max_steps=6),
        serve.ReactorRequest("ign512", 512, seed=3, t_end=2e-5),
        serve.ReactorRequest("ign256", 256, seed=4, t_end=2e-5),
    ]


def test_served_results_equal_alone_and_in_mix(torch_cuda):
    reqs = _mix()
    fps = [r.footprint() for r in reqs]
    budget = int(max(fps) * 1.6)     # forces queueing on memory as well as on slots
    srv = serve.Server(hbm_budget=budget, slots=4, chunk=4, keep_results=True)
    recs, wall = srv.run(reqs, real_time=False)
    assert len(recs) == len(reqs)
    assert srv.peak_used <= budget and srv.peak_active >= 2
    mix = srv.results
    for r in reqs:
        alone = serve.Server(hbm_budget=budget, slots=1, chunk=3, keep_results=True)
        alone.run([r], real_time=False)
        a, m = alone.results[r.name], mix[r.name]
        if isinstance(a, tuple):          # state and species of a reacting case
            assert all(torch.equal(x, y) for x, y in zip(a, m)), r.name
        else:
            assert torch.equal(a, m), r.name
    st = serve.latency_stats(recs)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_poisson_stream_runs(torch_cuda):
    reqs = serve.poisson_mix(6, rate=200.0, seed=99, kinds=("sod", "reactors"))
    srv = serve.Server(hbm_budget=8 << 30, slots=4, chunk=8)
    recs, wall = srv.run(reqs, real_time=True)
    assert len(recs) == 6
    for r in recs:
        assert r["done"] >= r["admit"] >= r["arrival"] - 1e-9
