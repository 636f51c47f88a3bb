"""The C++ host path: ocfd_run (host/ocfd_run.cpp over include/ocfd.hpp) drives
a level through the C ABI with no Python in the loop.  Its results are
checked against the oracle (Sod) and against the Python binding's run of the
same case (bit for bit: both call the same entry points in the same order)."""
import ctypes
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "opencfd-amr_amd" / "ocfd_run"


def run_driver(tmp_path, case, n, max_grid, steps, scheme, flags):
    out = tmp_path / f"{case}.bin"
    print("ocfd_run", case, n, max_grid, steps, scheme, flags, flush=True)
    r = subprocess.run([str(DRIVER), case, *(str(v) for v in n), str(max_grid), str(steps), str(scheme), str(flags),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    raw = out.read_bytes()
    nb = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    bx = np.frombuffer(raw[4:4 + 24 * nb], dtype=np.int32).reshape(nb, 6)
    data = np.frombuffer(raw[4 + 24 * nb:], dtype=np.float64)
    boxes, fabs, off = [], [], 0
    for b in range(nb):
        lo, hi = tuple(int(v) for v in bx[b, :3]), tuple(int(v) for v in bx[b, 3:])
        shp = (5, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        cnt = int(np.prod(shp))
        fabs.append(data[off:off + cnt].reshape(shp))
        off += cnt
        boxes.append((lo, hi))
    assert off == data.size
    return info, boxes, fabs


def test_driver_sod_matches_oracle(torch_cuda, tmp_path):
    import oracle as orc
    from ocfd import cases
    info, boxes, fabs = run_driver(tmp_path, "sod", (64, 8, 8), 32, 5, 2, 0)
    assert info["steps"] == 5
    want, g, kind = cases.sod(64, 8, 8, max_grid=(32, 32, 32))
    assert boxes == want
    og = gu.to_ogeom(g)
    fa, fb, fc = (orc.new_level_fabs(boxes) for _ in range(3))
    orc.init_case(kind, boxes, fa, og, orc.parm())
    orc.run(boxes, fa, fb, fc, og, orc.parm(), 0.5, 2, 0, 0, -1.0, 5)
    for b in range(len(boxes)):
        ref = fa[b].a[:, 4:-4, 4:-4, 4:-4]
        scale = np.abs(ref).reshape(5, -1).max(axis=1).reshape(5, 1, 1, 1) + 1e-300
        assert (np.abs(fabs[b] - ref) / scale).max() < 1e-10


@pytest.mark.parametrize("case,n,mg,steps,scheme,flags", [
    ("kh", (32, 32, 32), 16, 3, 2, 1 | 4),
    ("dmr", (128, 32, 2), 32, 6, 1, 0),
])
def test_driver_matches_python_binding(torch_cuda, tmp_path, case, n, mg, steps, scheme, flags):
    import ocfd
    from ocfd import cases
    info, boxes, fabs = run_driver(tmp_path, case, n, mg, steps, scheme, flags)
    if case == "kh":
        want, g, kind = cases.kh3d(n[0], max_grid=(mg, mg, mg))
    else:
        want, g, kind = cases.dmr(*n, max_grid=(mg, mg, mg))
    assert boxes == want
    lev = ocfd.Level(want, g)
    p = ocfd.default_parm()
    A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
    lev.init_state(kind, A, p, seed=42)
    st, t, dt = ocfd.run(lev, A, B, C, p, 0.5, scheme, flags, nsteps=steps)
    assert st == info["steps"] and t == info["time"] and dt == info["dt"]
    got = gu.download(lev, A)
    for b in range(len(want)):
        assert np.array_equal(got[b][:, 4:-4, 4:-4, 4:-4], fabs[b])
