"""Benchmark of the MI355X-native per-box advance.

Workload: 3D viscous Kelvin-Helmholtz shear layer, periodic 256^3 single
level, SSP-RK3, Sutherland viscosity, Pr 0.72 -- one step = one full RK3 time
step (3 ghost fills, 9 flux sweeps with the fused stage updates, one CFL
reduction) over all 256^3 cells, state resident in HBM.  The headline value is
the FP32-flux / FP64-state form; the plain FP64 form is measured beside it.

Prints one JSON line (rank 0, N=1).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "opencfd-amr_amd"))

# One hardware queue per serving slot (8 streams): with HIP's default of 4,
# two slots share a queue and a small case waits behind the other slot's
# kernels (served p50 77-86 ms -> 62-64 ms, p99 657-672 ms -> 458-511 ms with 8).
# Read by the HIP runtime when it starts, so it is set before torch loads
# (over an inherited value); the single-stream legs are unaffected.
os.environ["GPU_MAX_HW_QUEUES"] = "8"
# the CPU baselines' OpenMP threads stay on neighbouring cores (set before the
# oracle's OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per side of the KH cube")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128, help="cells per side of the CPU baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--no-chem", action="store_true", help="skip the 0D ignition batch")
    ap.add_argument("--reactors", type=int, default=65536)
    ap.add_argument("--no-serve", action="store_true", help="skip the served mixed stream")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--no-amr", action="store_true", help="skip the two-level DMR")
    ap.add_argument("--amr-only", action="store_true", help="only the two-level DMR (for profiling)")
    ap.add_argument("--amr-steps", type=int, default=10)
    ap.add_argument("--dmr-to-t", action="store_true",
                    help="also run the AMR double Mach reflection to t = 0.2 in both precision forms")
    ap.add_argument("--chem-only", action="store_true", help="only the 0D ignition batch, no CPU legs (for profiling)")
    ap.add_argument("--no-channel", action="store_true", help="skip the reacting channel")
    ap.add_argument("--channel-only", action="store_true", help="only the reacting channel (for profiling)")
    ap.add_argument("--channel-steps", type=int, default=20)
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 256 MiB device copy three times (HBM counter calibration)")
    return ap.parse_args()


def calib_copy():
    """Plain device-to-device copy of a known size: FETCH_SIZE / WRITE_SIZE of
    this kernel divided by 256 MiB gives the counter's scale factor."""
    import torch
    n = 256 * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": n * 8, "repeats": 3}))


def timed_run(lev, A, B, C, p, flags, steps, warmup):
    """Host-driven steps (the reference driver's form: the CFL minimum goes to
    the host every step)."""
    import torch
    import ocfd
    cfl = 0.5
    _, t, dt = ocfd.run(lev, A, B, C, p, cfl, 2, flags, nsteps=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ocfd.run(lev, A, B, C, p, cfl, 2, flags, nsteps=steps, time=t, dt=dt)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def timed_graph(lev, A, B, C, p, flags, steps, warmup):
    """Device-driven steps replayed from hipGraphs (dt rule on the device, no
    host round trip): `warmup` one-step replays, then one replay of a
    `steps`-step graph is timed on its stream."""
    import torch
    import ocfd
    s = torch.cuda.Stream()
    ctl = ocfd.new_ctl()
    with torch.cuda.stream(s):
        lev.ctl_start(A, ctl, p, 0.5, stream=s)
        g1 = lev.capture_steps(A, B, C, ctl, p, 1, 2, flags, stream=s)
        gk = lev.capture_steps(A, B, C, ctl, p, steps, 2, flags, stream=s)
    for _ in range(warmup):
        g1.launch(s)
    s.synchronize()
    t0 = time.perf_counter()
    gk.launch(s)
    s.synchronize()
    el = time.perf_counter() - t0
    c = ocfd.read_ctl(ctl)
    assert c.steps == warmup + steps and not c.done, (c.steps, c.done)
    return el


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_kh3d_256_fp32.txt"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration)."""
    if not TRAFFIC_FILE.exists():
        return None
    txt = "".join(l for l in TRAFFIC_FILE.read_text().splitlines(True) if not l.startswith("#"))
    k = json.loads(txt)["kernels"].get(kernel)
    return k["traffic_bytes"] if k else None


# Compulsory HBM bytes per cell of one SSP-RK3 stage (FP64 state, 5 comps):
# stage 1 reads U^n and writes U^1 (80 B); stages 2 and 3 read U^(k) and U^n
# and write U^(k+1) (120 B each): 320 B per cell-update.
COMPULSORY_STAGE_BYTES = (80.0, 120.0, 120.0)


def kernel_roofline(lev, A, B, C, p, flags, ncell):
    """Per-kernel average durations of one RK stage (stage-2 form, HIP events on
    the launch stream) and the roofline figures of the dominant kernel.

    `achieved` counts the compulsory bytes of the launch: the z sweep is the
    kernel that reads U^(k) and U^n and writes U^(k+1), so one launch of it
    (stage-2 form) moves 120 B per cell of the level by necessity; the sweeps
    before it move only intermediates (primitives, partial divergences) that a
    fused kernel would not need, so their compulsory share is 0.  The bytes
    the kernels actually read and write by design are reported beside it
    (`design_bytes`, the per-kernel `algorithmic_bytes`)."""
    lev.fill_boundary(A, 0.0)
    ms = lev.profile_stage(A, B, C, 0.25, 0.75, 1e-9, p, flags, reps=5)
    names = ["sweep_x_primitives", "sweep_y", "sweep_z_update", "fill_boundary"]
    # algorithmic HBM bytes of each kernel of a stage; f = bytes of the flux
    # type (4 FP32, 8 FP64), state FP64 with 5 comps:
    #   sweep_x_primitives: on grow(bx,3) conserved in (40 B) and 6 primitives
    #                   out; per valid cell the 5-comp partial divergence out
    #   sweep_y: primitives in, partial in and out                     = 16 f
    #   sweep_z_update: primitives, partial, U^(k) and U^n in (80 B),
    #                   U^(k+1) out (40 B)                             = 11 f + 120
    #   fill: ghost cells only, read+write 80 B per ghost cell
    nx = lev.boxes[0][1][0] - lev.boxes[0][0][0] + 1
    ghost = (nx + 8) ** 3 - nx ** 3
    f = 4.0 if flags & 4 else 8.0
    bytes_ = [(40.0 + 6 * f) * (nx + 6) ** 3 + 5 * f * ncell, 16 * f * ncell, (11 * f + 120.0) * ncell,
              80.0 * ghost]
    kern = {}
    for nm, t, b in zip(names, ms, bytes_):
        kern[nm] = {"ms": t, "GBps": b / (t * 1e-3) / 1e9, "algorithmic_bytes": b}
    dom = max(range(3), key=lambda i: ms[i])
    compulsory = [0.0, 0.0, COMPULSORY_STAGE_BYTES[1] * ncell]
    achieved = compulsory[dom] / (ms[dom] * 1e-3) / 1e9
    return names[dom], achieved, compulsory[dom], kern


def chemistry_0d(n_react, cpu_sample=256, cpu_sample_fixed=2048, cpu_sample_ros=2048, cpu=True):
    """0D constant-volume H2/air ignition: n_react seeded reactors (T0 900-1500 K,
    p0 1-10 atm, phi 0.5-2) integrated to 1 ms with per-reactor adaptive
    substeps in one persistent launch: RODAS3 (the library's stiff
    integrator) and, beside it, step doubling on the reference's
    linearised-implicit step; the oracle's identical integrators are timed on
    samples of the same reactors."""
    import numpy as np
    import torch
    import ocfd
    T0, p0, phi = ocfd.ignition_batch(n_react, seed=7)

    def timed(method):
        R = ocfd.Reactors(T0, p0, phi)
        getattr(R, method)(1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)   # warm-up (also first-touch of the code)
        R = ocfd.Reactors(T0, p0, phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        getattr(R, method)(1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        steps = R.steps.cpu().numpy()
        rej = R.rejects.cpu().numpy()
        T = R.temperature().cpu().numpy()
        return el, steps, rej, T

    el_r, steps_r, rej_r, T_r = timed("advance_rosenbrock")
    el, steps, rej, T = timed("advance_adaptive")
    out = {"reactors": n_react, "t_end_s": 1e-3, "rtol": 1e-4, "atol": 1e-9,
           "integrator": "RODAS3 (4-stage Rosenbrock, order 3, full Jacobian)",
           "ms": el_r * 1e3, "reactors_per_s": n_react / el_r,
           "accepted_steps_mean": float(steps_r.mean()), "rejected_steps_mean": float(rej_r.mean()),
           "failed": int((steps_r < 0).sum()), "ignited_frac": float((T_r > 1.5 * T0).mean()),
           "T_finite": bool(np.isfinite(T_r).all()),
           "T_rel_diff_vs_step_doubling": {
               "median": float(np.median(np.abs(T_r - T) / T)),
               "p99": float(np.percentile(np.abs(T_r - T) / T, 99)),
               "above_1e-3": int((np.abs(T_r - T) / T > 1e-3).sum()),
               "note": "the reactors above 1e-3 are mid-ignition at 1 ms, where a small change of the ignition "
                       "delay moves T steeply; on such reactors RODAS3 is the closer of the two to a tight "
                       "Radau solution"},
           "step_doubling": {"ms": el * 1e3, "reactors_per_s": n_react / el,
                             "substeps_per_s": float((steps.clip(0) + rej).sum() * 3 / el),
                             "accepted_steps_mean": float(steps.mean()), "failed": int((steps < 0).sum()),
                             "ignited_frac": float((T > 1.5 * T0).mean())}}
    if not cpu:
        return out
    # the reference's own scheme: repeated linearised-implicit chemical_advance
    # steps at a fixed dt (Source/CHEM/ChemAdvance.cpp:7-85, the fixed-step
    # driver of Test/chemical_test.cpp:947-1006), 1000 steps of 1 us to 1 ms
    dt_fix, n_fix = 1e-6, 1000
    Rf = ocfd.Reactors(T0, p0, phi)
    Rf.advance_fixed(dt_fix, 10)
    Rf = ocfd.Reactors(T0, p0, phi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Rf.advance_fixed(dt_fix, n_fix)
    torch.cuda.synchronize()
    el_f = time.perf_counter() - t0
    out["fixed_step"] = {"dt_s": dt_fix, "steps": n_fix, "ms": el_f * 1e3, "reactors_per_s": n_react / el_f,
                         "reactor_steps_per_s": n_react * n_fix / el_f}
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    ncpu, model = host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS", ncpu))

    def adaptive():
        u, r = orc.reactor_init(T0[:cpu_sample], p0[:cpu_sample], phi[:cpu_sample])
        orc.reactor_adaptive(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)

    def fixed():
        u, r = orc.reactor_init(T0[:cpu_sample_fixed], p0[:cpu_sample_fixed], phi[:cpu_sample_fixed])
        orc.reactor_fixed(u, r, dt_fix, n_fix)
    def rosenbrock():
        u, r = orc.reactor_init(T0[:cpu_sample_ros], p0[:cpu_sample_ros], phi[:cpu_sample_ros])
        orc.reactor_rosenbrock(u, r, 1e-3, h0=1e-8, rtol=1e-4, atol=1e-9)
    med_r, ts_r = median_of(rosenbrock)
    med_a, ts_a = median_of(adaptive)
    med_f, ts_f = median_of(fixed)
    out["cpu_baseline"] = {"value": cpu_sample_ros / med_r, "unit": "reactors/s", "cores": cores, "kind": "port",
                           "sample": f"oracle RODAS3 integrator, first {cpu_sample_ros} reactors, median of 5 "
                                     f"after 1 warm-up ({', '.join(f'{t:.2f}' for t in ts_r)} s)",
                           "nproc": ncpu, "cpu_model": model}
    out["step_doubling"]["cpu_baseline"] = {
        "value": cpu_sample / med_a, "unit": "reactors/s", "cores": cores, "kind": "port",
        "sample": f"oracle step-doubling integrator, first {cpu_sample} reactors, median of 5 after 1 warm-up "
                  f"({', '.join(f'{t:.2f}' for t in ts_a)} s)"}
    out["fixed_step"]["cpu_baseline"] = {
        "value": cpu_sample_fixed / med_f, "unit": "reactors/s", "cores": cores, "kind": "port",
        "sample": f"oracle fixed-step chemical_advance (bitwise equal to the reference's kernels), first "
                  f"{cpu_sample_fixed} reactors, {n_fix} steps of {dt_fix:g} s, median of 5 after 1 warm-up "
                  f"({', '.join(f'{t:.2f}' for t in ts_f)} s)"}
    return out


def serving(nreq, rate, seed=99, reactor_method="rosenbrock"):
    """Served mixed stream: Poisson arrivals (seeded) of 2D Sod/DMR slabs,
    3D periodic KH boxes, reacting EB channels and 0D ignition batches (their
    integrator: reactor_method) through the continuous-batching request loop;
    latency = completion - arrival."""
    import numpy as np
    from ocfd import serve

    def mix(n, r, sd):
        out = serve.poisson_mix(n, rate=r, seed=sd)
        for q in out:
            if isinstance(q, serve.ReactorRequest):
                q.method = reactor_method
        return out
    warm = mix(6, 1e3, seed + 1)
    serve.Server(hbm_budget=32 << 30, slots=8, chunk=8).run(warm, real_time=False)   # first-use costs
    reqs = mix(nreq, rate, seed)
    srv = serve.Server(hbm_budget=32 << 30, slots=8, chunk=8)
    recs, wall = srv.run(reqs, real_time=True)
    st = serve.latency_stats(recs)
    cell_updates = sum(r["cells"] * (r["steps"] or 0) for r in recs)
    kinds = {}
    for r in recs:
        k = r["name"].rstrip("0123456789")
        kinds.setdefault(k, []).append(r["latency"] * 1e3)
    return {"requests": len(recs), "rate_per_s": rate, "wall_s": wall, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "mean_ms": st["mean_ms"], "requests_per_s": len(recs) / wall,
            "flow_Mcell_updates_per_s": cell_updates / wall / 1e6,
            "p50_ms_by_kind": {k: float(np.percentile(v, 50)) for k, v in kinds.items()},
            "hbm_budget_GiB": 32, "slots": 8, "peak_resident": srv.peak_active,
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))}


def chem_channel(steps=20, warmup=5, nx=512, ny=128, max_grid=64, cpu_sample=True):
    """The reacting H2/air channel past a cylinder (nx x ny x 2 cells,
    max_grid boxes, regular boxes through the multispecies WENO path, cut
    boxes through the EB-CHEM path): Strang-split RK2 steps at CFL 20 x 1e-8 s,
    device-driven and replayed from a hipGraph of `steps` steps.  The oracle's
    identical step is timed on the 128 x 32 x 2 channel for the CPU line."""
    import torch
    import ocfd
    from ocfd import cases
    boxes, g, kind, sp, (state, spec) = cases.chem_channel(nx, ny, 2, max_grid=max_grid)
    lev = ocfd.Level(boxes, g, eb=sp)
    lev.set_inflow_species(spec)
    A, As, B, Bs = lev.new_mf(), lev.new_spec_mf(), lev.new_mf(), lev.new_spec_mf()
    p = ocfd.default_parm()
    cfl = 20.0
    lev.init_species_uniform(state, spec, A, As)
    s = torch.cuda.Stream()
    ctl = ocfd.new_ctl()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        lev.ctl_start_chem(A, ctl, p, cfl, stream=s)
        Gw = lev.capture_steps_chem(A, As, B, Bs, ctl, p, warmup, scheme=1, flags=ocfd.VISC, stream=s)
        G = lev.capture_steps_chem(A, As, B, Bs, ctl, p, steps, scheme=1, flags=ocfd.VISC, stream=s)
        Gw.launch(s)
    s.synchronize()
    t0 = time.perf_counter()
    G.launch(s)
    s.synchronize()
    el = time.perf_counter() - t0
    c = ocfd.read_ctl(ctl)
    del G, Gw
    ncell = nx * ny * 2
    kinds = lev.box_types()
    out = {"workload": f"reacting H2/air channel past a cylinder, {nx} x {ny} x 2 cells, {len(boxes)} boxes "
                       f"({kinds.count(0)} regular, {kinds.count(1)} cut, {kinds.count(2)} covered), "
                       "Strang-split RK2 + linearised-implicit chemistry, dt = 20 x 1e-8 s",
           "steps": steps, "ms_per_step": 1e3 * el / steps, "Mcell_updates_per_s": ncell * steps / el / 1e6,
           "status": c.status, "steps_done": c.steps, "state_finite": bool(torch.isfinite(A).all()),
           "form": "device-driven steps replayed from a hipGraph"}
    if cpu_sample:
        sys.path.insert(0, str(ROOT / "oracle"))
        import ctypes
        import oracle as orc
        cb, cg, _, csp, (cst, csn) = cases.chem_channel(128, 32, 2, max_grid=16)
        og = orc.OGeom()
        ctypes.memmove(ctypes.byref(og), ctypes.byref(cg), ctypes.sizeof(og))
        ebs = []
        for lo, hi in cb:
            glo, ghi = orc.grown(lo, hi)
            gm = ocfd.eb_geometry(csp, cg, lo, hi, 4)
            ebs.append(orc.EbFab(glo, ghi, gm["flag"], gm["vfrac"], gm["apx"], gm["apy"], gm["apz"]))
        fa, fs = orc.new_level_fabs(cb), orc.new_spec_fabs(cb)
        fb, fbs = orc.new_level_fabs(cb), orc.new_spec_fabs(cb)
        for f, fsp in zip(fa, fs):
            f.a[:] = cst[:, None, None, None]
            fsp.a[:] = csn[:, None, None, None]
        orc.set_eb(ebs)
        orc.set_inflow_species(csn)
        dt = cfl * ocfd.CHEM_ESTDT
        state = {"t": 0.0}

        nsamp = 10

        def ten_steps():
            for _ in range(nsamp):
                assert orc.advance_chem(cb, fa, fs, fb, fbs, og, orc.parm(), state["t"], dt, 0, 1, 1, 1) == 0
                state["t"] += dt
        try:
            med, ts = median_of(ten_steps)
        finally:
            orc.set_eb(None)
            orc.set_inflow_species(None)
        ncpu, model = host_cpu()
        out["cpu_baseline"] = {"value": 128 * 32 * 2 * nsamp / med / 1e6, "unit": "Mcell-updates/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle Strang-split steps of the 128 x 32 x 2 channel (16 x 16 boxes), "
                                         f"{nsamp} steps per run, "
                                         f"median of 5 after 1 warm-up ({', '.join(f'{t:.2f}' for t in ts)} s)",
                               "nproc": ncpu, "cpu_model": model}
    return out


def _cells(boxes):
    n = 0
    for lo, hi in boxes:
        n += (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
    return n


def amr_dmr(steps, warmup=40, nx=1024, ny=256, max_level=2, cpu_sample=True, uniform=True, fp32=False):
    """Double Mach reflection with AMR: base 1024x256x2 in uneven boxes of 32 to
    128 cells (seeded), max_level 2 at ratio 2, tagged on density gradient
    (refine_dengrad 0.05), regrid every 2 steps of each level, blocking factor 2,
    max_grid_size 64, AMReX's default grid efficiency 0.7; SSP-RK3, CFL 0.5,
    FP64, planar stepping of the z-uniform flow.  A cell update is one cell
    advanced by its level's dt (a level-l cell counts 2^l times per coarse
    step).  The same time span on the uniform grid at the finest resolution is
    timed beside it.  fp32: both runs in the FP32-flux / FP64-state form."""
    import torch
    import ocfd
    from ocfd import amr, cases
    p = ocfd.default_parm()
    boxes, g, kind = cases.dmr(nx, ny, 2, sizes=(32, 64, 128), seed=1234)
    kw = dict(scheme=2, max_level=max_level, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, grid_eff=0.7,
              max_grid_size=64)
    flags = ocfd.PLANAR | (ocfd.FLUX_FP32 if fp32 else 0)
    H = amr.Hierarchy(boxes, g, p, 0.5, kind, flags=flags, **kw)
    for _ in range(warmup):
        H.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    upd, upd2, dts, per_level, nbox = 0, 0, [], [0.0] * (max_level + 1), [0.0] * (max_level + 1)
    for _ in range(steps):
        dts.append(H.step())
        c = H.cells()
        for l, n in enumerate(c):
            upd += n * 2 ** l
            upd2 += n // (2 * 2 ** l) * 2 ** l     # 2D cells: the level has 2 * 2^l z layers
            per_level[l] += n / steps
            nbox[l] += H.levels[l].nbox / steps
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nc = _cells(boxes)
    form = "FP32 fluxes / FP64 state" if fp32 else "FP64"
    out = {"workload": f"DMR {nx}x{ny}x2 base (uneven boxes 32-128) + {max_level} levels of ratio 2, SSP-RK3, {form}",
           "coarse_steps": steps, "after_steps": warmup, "ms_per_coarse_step": 1e3 * el / steps,
           "cell_updates_per_s": upd / el, "cell_updates_2d_per_s": upd2 / el,
           "note": "cell_updates counts every z layer (2, 4, 8 per level) as the reference's 3D build does; the "
                   "planar levels sweep their first z layer and copy it, so cell_updates_2d (one layer) is the "
                   "work-honest rate of this 2D flow",
           "cells_per_level": per_level, "boxes_per_level": nbox,
           "fraction_of_domain_per_level": [per_level[l] / (nc * 8 ** l) for l in range(max_level + 1)]}
    del H
    torch.cuda.empty_cache()
    if uniform:
        r = 2 ** max_level
        fb, fg, _ = cases.dmr(r * nx, r * ny, 2 * r, max_grid=(128, 128, 2 * r))
        lev = ocfd.Level(fb, fg).set_planar(True)
        A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
        lev.init_state(kind, A, p)
        lev.advance(A, B, C, 0.0, dts[0] / r, 2, p, flags)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tt = 0.0
        for dt in dts:
            for _ in range(r):
                lev.advance(A, B, C, tt, dt / r, 2, p, flags)
                tt += dt / r
        torch.cuda.synchronize()
        elu = time.perf_counter() - t0
        out["uniform_finest"] = f"{r * nx}x{r * ny}x{2 * r}, {r} steps per coarse step"
        out["uniform_finest_ms_per_coarse_step"] = 1e3 * elu / steps
        out["speedup_vs_uniform_finest"] = elu / el
        del lev, A, B, C
        torch.cuda.empty_cache()
    if cpu_sample:
        sys.path.insert(0, str(ROOT / "oracle"))
        import ctypes
        import oracle as orc
        import amr_oracle as ao
        sb, sg, _ = cases.dmr(256, 64, 2, max_grid=(64, 64, 2))
        og = orc.OGeom()
        ctypes.memmove(ctypes.byref(og), ctypes.byref(sg), ctypes.sizeof(og))
        O = ao.Hierarchy(sb, og, orc.parm(), 0.5, kind, **kw)
        t0 = time.perf_counter()
        upd_c = 0
        for _ in range(2):
            O.step()
            upd_c += sum(_cells(O.boxes[l]) * 2 ** l for l in range(O.finest + 1))
        el_c = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": upd_c / el_c, "unit": "cell-updates/s", "cores": cores, "kind": "port",
                               "sample": f"AMR oracle (C++ fluxes, numpy AMR steps), DMR 256x64x2 + {max_level} "
                                         f"levels, 2 coarse steps ({el_c:.1f} s)"}
    return out


def amr_dmr_to_time(stop_time=0.2, nx=1024, ny=256, max_level=2):
    """The DMR configuration run to t = 0.2 in the FP64 form and in the
    FP32-flux / FP64-state form: wall time, coarse steps, and the relative L1
    density difference between the two forms on the base level (averaged down
    from the finer levels)."""
    import torch
    import ocfd
    from ocfd import amr, cases
    p = ocfd.default_parm()
    boxes, g, kind = cases.dmr(nx, ny, 2, sizes=(32, 64, 128), seed=1234)
    kw = dict(scheme=2, max_level=max_level, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, grid_eff=0.7,
              max_grid_size=64, stop_time=stop_time)
    res, rho = {}, {}
    for name, flags in (("fp64", ocfd.PLANAR), ("fp32_flux", ocfd.PLANAR | ocfd.FLUX_FP32)):
        H = amr.Hierarchy(boxes, g, p, 0.5, kind, flags=flags, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 0
        while H.time < stop_time:
            H.step()
            n += 1
            if n % 200 == 0:
                print(f"dmr {name}: step {n} t={H.time:.4f} cells per level {H.cells()}", file=sys.stderr,
                      flush=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[name] = {"wall_s": el, "coarse_steps": n, "time": H.time, "cells_per_level": H.cells()}
        rho[name] = torch.cat([H.levels[0].fab_view(H.S[0], b)[0, 4:-4, 4:-4, 4:-4].reshape(-1)
                               for b in range(H.levels[0].nbox)]).cpu()
        del H
        torch.cuda.empty_cache()
    d = (rho["fp32_flux"] - rho["fp64"]).abs().sum() / rho["fp64"].abs().sum()
    res["rel_L1_density_fp32_vs_fp64"] = float(d)
    return res


def _affinity():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# taken at import, before an OpenMP runtime binds this thread to one core
HOST_NCPU = _affinity()


def host_cpu():
    """Host core count (the process's CPU set when it started) and CPU model."""
    ncpu = HOST_NCPU
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return ncpu, model


def median_of(fn, runs=5, warmup=1):
    """Wall seconds of fn(): median of `runs` timed runs after `warmup` untimed."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], ts


def cpu_baseline(n, steps):
    """The FP64 C++/OpenMP oracle (the reference's per-box path restated) on
    the host cores: KH n^3 periodic viscous, `steps` SSP-RK3 steps from the
    seeded initial data; median of 5 runs after one warm-up."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ctypes
    import oracle as orc
    from ocfd import cases
    boxes, g, kind = cases.kh3d(n)
    og = orc.OGeom()
    ctypes.memmove(ctypes.byref(og), ctypes.byref(g), ctypes.sizeof(og))
    A, B, C = (orc.new_level_fabs(boxes) for _ in range(3))

    def one():
        orc.init_case(kind, boxes, A, og, orc.parm(), seed=42)
        st, _, _ = orc.run(boxes, A, B, C, og, orc.parm(), 0.5, 2, 1, 0, -1.0, steps)
        assert st == steps
    med, ts = median_of(one)
    ncpu, model = host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS", ncpu))
    return {"value": n ** 3 * steps / med / 1e6, "unit": "Mcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 C++/OpenMP oracle, KH {n}^3 periodic viscous, {steps} SSP-RK3 steps; median of 5 "
                      f"runs after 1 warm-up ({', '.join(f'{t:.2f}' for t in ts)} s)",
            "nproc": ncpu, "cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}


def fp32_error(lev, A, B, C, p, kind, steps=100):
    """FP32-flux form against the FP64 form on the same initial data, `steps`
    SSP-RK3 steps at one fixed dt (CFL 0.5 of the initial state): relative
    (L1, Linf) per conserved variable, and whether every value is finite."""
    import torch
    import ocfd
    lev.init_state(kind, A, p, seed=42)
    A32 = A.clone()
    dt = lev.estdt(A, 0.5, p)
    for S, flags in ((A, ocfd.VISC), (A32, ocfd.VISC | ocfd.FLUX_FP32)):
        t = 0.0
        for _ in range(steps):
            lev.advance(S, B, C, t, dt, 2, p, flags)
            t += dt
    torch.cuda.synchronize()
    e = ocfd.level_error(lev, A32, A)
    finite = bool(torch.isfinite(A32).all())
    del A32
    names = ["rho", "mx", "my", "mz", "E"]
    return {"steps": steps, "dt": dt, "finite": finite, "rel_L1": dict(zip(names, (x[0] for x in e))),
            "rel_Linf": dict(zip(names, (x[1] for x in e))), "target_rel_L1_rho": 1e-4}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding is a "
              "design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    import torch
    import ocfd
    from ocfd import cases

    torch.cuda.set_device(0)
    if args.calib_copy:
        calib_copy()
        return
    if args.chem_only:
        print(json.dumps({"chemistry_0d": chemistry_0d(args.reactors, cpu=False)}))
        return
    if args.channel_only:
        print(json.dumps({"chem_channel": chem_channel(args.channel_steps, cpu_sample=False)}))
        return
    if args.amr_only:
        out = {"amr_dmr": amr_dmr(args.amr_steps, cpu_sample=False),
               "amr_dmr_fp32_flux": amr_dmr(args.amr_steps, cpu_sample=False, fp32=True)}
        if args.dmr_to_t:
            out["amr_dmr_to_t02"] = amr_dmr_to_time()
        print(json.dumps(out))
        return
    n = args.n
    boxes, g, kind = cases.kh3d(n)
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    ncell = n ** 3

    flags32 = ocfd.VISC | ocfd.FLUX_FP32
    lev.init_state(kind, A, p, seed=42)
    lev.fill_boundary(A, 0.0)
    ch32, tune32 = lev.autotune(A, p, flags32, reps=3)
    lev.init_state(kind, A, p, seed=42)
    t32 = timed_graph(lev, A, B, C, p, flags32, args.steps, args.warmup)
    finite32 = bool(torch.isfinite(A).all())
    lev.init_state(kind, A, p, seed=42)
    t32h = timed_run(lev, A, B, C, p, flags32, args.steps, args.warmup)
    dom, achieved, compulsory, kern32 = kernel_roofline(lev, A, B, C, p, flags32, ncell)
    value = ncell * args.steps / t32 / 1e6
    err32 = fp32_error(lev, A, B, C, p, kind) if not args.skip_fp64 else None

    res64 = None
    if not args.skip_fp64:
        flags64 = ocfd.VISC
        lev.init_state(kind, A, p, seed=42)
        lev.fill_boundary(A, 0.0)
        ch64, tune64 = lev.autotune(A, p, flags64, reps=3)
        lev.init_state(kind, A, p, seed=42)
        t64 = timed_run(lev, A, B, C, p, flags64, args.steps, args.warmup)
        lev.init_state(kind, A, p, seed=42)
        t64g = timed_graph(lev, A, B, C, p, flags64, args.steps, args.warmup)
        _, _, _, kern64 = kernel_roofline(lev, A, B, C, p, flags64, ncell)
        res64 = {"value": ncell * args.steps / t64 / 1e6, "ms_per_step": 1e3 * t64 / args.steps,
                 "autotune": {"march_chunk": ch64, "y_plus_z_sweep_ms": tune64},
                 "form": "host-driven steps (plain FP64 GPU form)",
                 "graph_value": ncell * args.steps / t64g / 1e6, "kernels": kern64}
    del A, B, C, lev
    torch.cuda.empty_cache()

    amr_res = None
    if not args.no_amr:
        amr_res = amr_dmr(args.amr_steps, cpu_sample=not args.no_cpu_baseline)
        amr_res["fp32_flux"] = amr_dmr(args.amr_steps, cpu_sample=False, fp32=True)
        if args.dmr_to_t:
            amr_res["to_t02"] = amr_dmr_to_time()

    serve_res = None
    if not args.no_serve:
        serve_res = serving(args.serve_requests, args.serve_rate)
        sd = serving(args.serve_requests, args.serve_rate, reactor_method="adaptive")
        serve_res["step_doubling_ignition"] = {"p50_ms_by_kind": sd["p50_ms_by_kind"], "p50_ms": sd["p50_ms"],
                                               "p99_ms": sd["p99_ms"]}

    channel = None
    if not args.no_channel:
        channel = chem_channel(args.channel_steps, cpu_sample=not args.no_cpu_baseline)

    chem = None
    if not args.no_chem:
        chem = chemistry_0d(args.reactors)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps)

    out = {
        "metric": "Mcell-updates/s + % HBM peak; p50/p99 case latency, FP32-flux vs FP64",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state",
        "data": "synthetic (seeded KH initial data)",
        "config": {"workload": f"3D viscous Kelvin-Helmholtz, periodic {n}^3 single level, SSP-RK3, "
                               "FP32 fluxes / FP64 state", "cells": ncell, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes": compulsory,
                     "algorithmic_bytes_basis": "compulsory: U^(k), U^n read + U^(k+1) written, 120 B per cell "
                                                "(stage-2 form) x 256^3 cells",
                     "design_bytes": kern32[dom]["algorithmic_bytes"],
                     "design_frac": kern32[dom]["GBps"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom) if n == 256 else None,
                     "traffic_source": "profiles/hbm_traffic_kh3d_256_fp32.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"},
        "hbm_frac_step": sum(COMPULSORY_STAGE_BYTES) * value * 1e6 / 1e9 / HBM_PEAK_GBS,
        "state_finite": finite32,
        "fp32_vs_fp64": err32,
        "kernels": kern32,
        "form": "FP32-flux device-driven steps replayed from a hipGraph",
        "autotune": {"march_chunk": ch32, "y_plus_z_sweep_ms": tune32},
        "host_driven_fp32": {"value": ncell * args.steps / t32h / 1e6, "ms_per_step": 1e3 * t32h / args.steps},
        "fp64_form": res64,
        "fp32_over_fp64": (value / res64["value"]) if res64 else None,
        "serving": serve_res,
        "amr_dmr": amr_res,
        "cpu_baseline": cpu,
        "chemistry_0d": chem,
        "chem_channel": channel,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
