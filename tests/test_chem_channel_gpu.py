"""The reacting H2/air channel past a cylinder (cases.chem_channel): its
Strang-split multispecies steps against the oracle on a 128x32 channel
(regular boxes through the WENO multispecies path, cut boxes through the
EB-CHEM path, the constant premixed inflow at x-lo), the device-driven and
graph-replayed steps bit-identical to the host loop, and the serving
footprint covering what the case holds on the device."""
import numpy as np
import pytest

import gpu_util as gu
from test_chem_flow_gpu import _eb_fabs

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
CFL = 20.0            # Exec/9_CHEMEB/inputs: ebr.cfl = 20, so dt = 20 x 1e-8 s


def _channel(nx=128, ny=32, mg=16):
    from ocfd import cases
    boxes, g, kind, sp, (state, spec) = cases.chem_channel(nx, ny, 2, max_grid=mg)
    return boxes, g, sp, state, spec


def _level(boxes, g, sp, state, spec):
    import ocfd
    lev = ocfd.Level(boxes, g, eb=sp)
    lev.set_inflow_species(spec)
    S, P = lev.new_mf(), lev.new_spec_mf()
    lev.init_species_uniform(state, spec, S, P)
    return lev, S, P


def test_chem_channel_matches_oracle(torch_cuda, orc):
    import ocfd
    boxes, g, sp, state, spec = _channel()
    lev, S, P = _level(boxes, g, sp, state, spec)
    kinds = lev.box_types()
    assert 0 in kinds and 1 in kinds
    og = gu.to_ogeom(g)
    fa, fs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    for f, fsp in zip(fa, fs):
        f.a[:] = state[:, None, None, None]
        fsp.a[:] = spec[:, None, None, None]
    ebs = _eb_fabs(orc, boxes, g, sp)
    p = ocfd.default_parm()
    B, Bp = lev.new_mf(), lev.new_spec_mf()
    fb, fbs = orc.new_level_fabs(boxes), orc.new_spec_fabs(boxes)
    dt = CFL * ocfd.CHEM_ESTDT
    orc.set_eb(ebs)
    orc.set_inflow_species(spec)
    try:
        t = 0.0
        for _ in range(4):
            lev.advance_chem(S, P, B, Bp, t, dt, p, 0, 1, ocfd.VISC, True)
            assert orc.advance_chem(boxes, fa, fs, fb, fbs, og, orc.parm(), t, dt, 0, 1, 1, 1) == 0
            t += dt
    finally:
        orc.set_eb(None)
        orc.set_inflow_species(None)
    v = lambda arrs: [a[:, 4:-4, 4:-4, 4:-4] for a in arrs]
    e = gu.level_rel_err(v(gu.download(lev, S)), v([f.a for f in fa]))
    got_p = v([lev.fab_view(P, b, 4, 9).cpu().numpy() for b in range(lev.nbox)])
    ref_p = v([f.a for f in fs])
    ep = max(np.abs(a - r).max() for a, r in zip(got_p, ref_p)) / max(np.abs(r).max() for r in ref_p)
    print("chem channel 128x32, 4 steps: state (L1, Linf)", e, "species Linf", ep)
    assert max(max(x) for x in e) < TOL64
    assert ep < TOL64
    # the mixture reacted: radicals formed from the premixed H2/O2/N2
    assert max(np.abs(r[3]).max() for r in got_p) > 0          # H


def test_chem_channel_device_steps_equal_host_loop(torch_cuda):
    """Device-driven steps (dt and time on the device) and a replayed graph of
    them give the host loop's bits: state, species, time and step count."""
    import torch
    import ocfd
    boxes, g, sp, state, spec = _channel(64, 16, 16)
    flags = ocfd.VISC
    p = ocfd.default_parm()
    lev, S1, P1 = _level(boxes, g, sp, state, spec)
    B, Bp = lev.new_mf(), lev.new_spec_mf()
    steps, t_host, _ = ocfd.run_chem(lev, S1, P1, B, Bp, p, CFL, nsteps=6, scheme=1, flags=flags)
    S2, P2 = lev.new_mf(), lev.new_spec_mf()
    lev.init_species_uniform(state, spec, S2, P2)
    ctl = ocfd.new_ctl()
    lev.ctl_start_chem(S2, ctl, p, CFL, max_steps=6)
    for _ in range(8):                       # the last two are no-ops
        lev.step_device_chem(S2, P2, B, Bp, ctl, p, scheme=1, flags=flags)
    c = ocfd.read_ctl(ctl)
    assert c.steps == steps == 6 and c.done == 1 and c.status == 0 and c.time == t_host
    assert torch.equal(S1, S2) and torch.equal(P1, P2)
    # graph of 3 steps, replayed twice on its own stream
    S3, P3 = lev.new_mf(), lev.new_spec_mf()
    lev.init_species_uniform(state, spec, S3, P3)
    s = torch.cuda.Stream()
    ctl3 = ocfd.new_ctl()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        lev.ctl_start_chem(S3, ctl3, p, CFL, max_steps=6, stream=s)
        G = lev.capture_steps_chem(S3, P3, B, Bp, ctl3, p, 3, scheme=1, flags=flags, stream=s)
        G.launch(s)
        G.launch(s)
    s.synchronize()
    assert torch.equal(S3, S2) and torch.equal(P3, P2)
    del G


def test_chem_channel_footprint_covers_device_memory(torch_cuda):
    import ocfd
    from ocfd import cases, serve
    req = serve.FlowRequest("c", lambda: cases.chem_channel(128, 32, 2, max_grid=16), flags=ocfd.VISC, cfl=CFL,
                            scheme=1, max_steps=2)
    boxes, g, kind, sp, (state, spec) = req.made()
    lev, S, P = _level(boxes, g, sp, state, spec)
    B, Bp = lev.new_mf(), lev.new_spec_mf()
    lev.advance_chem(S, P, B, Bp, 0.0, CFL * ocfd.CHEM_ESTDT, ocfd.default_parm(), 0, 1, ocfd.VISC, True)
    actual = lev.device_bytes() + 2 * 8 * lev.mf_size() + 2 * 8 * lev.mf_size(ncomp=9) + ocfd.CTL_BYTES
    fp = req.footprint()
    print("chem channel footprint", fp, "held", actual)
    assert actual <= fp <= 2 * actual
