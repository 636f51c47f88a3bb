"""Plotfile format and fcompare, on the CPU.

The plotfile oracle (oracle/plotfile_oracle.py) is pinned by a write/read
round trip and by the closed-form plot variables of the double Mach
reflection's two states (post-shock p = 116.5, u = 8.25 cos 30 deg; pre-shock
p = 1 at rest, Exec/1_DMR/prob.H).  The library's fcompare
(ocfd_plotfile_compare, host code) is pinned against norms computed directly
from the data: identical files agree, a known perturbation gives the known
norm-0/1/2 absolute and relative errors, tolerances and NaNs are reported as
fcompare reports them, and differently boxed grids compare equal only with
allow_diff_grids."""
import ctypes
import re

import numpy as np
import pytest

import plotfile_oracle as po


def _dmr_level(orc, nx=32, ny=8, max_grid=(16, 8, 2)):
    from ocfd import cases
    boxes, g, kind = cases.dmr(nx, ny, 2, max_grid=max_grid)
    og = orc.OGeom()
    ctypes.memmove(ctypes.byref(og), ctypes.byref(g), ctypes.sizeof(og))
    A = orc.new_level_fabs(boxes)
    orc.init_case(kind, boxes, A, og, orc.parm())
    return boxes, A, og


def test_oracle_round_trip_and_dmr_states(orc, tmp_path):
    boxes, A, og = _dmr_level(orc)
    po.write(tmp_path / "plt", [(boxes, A, og)], 0.25, [7], 1.4)
    P = po.read(tmp_path / "plt")
    assert P["names"] == list(po.NAMES) and P["time"] == 0.25
    L = P["levels"][0]
    assert [tuple(map(tuple, b)) for b in L["boxes"]] == [tuple(map(tuple, b)) for b in boxes]
    for (lo, hi), fab, got in zip(boxes, A, L["fabs"]):
        assert np.array_equal(got, po.plot_vars(fab.valid(lo, hi), 1.4))
    allv = np.concatenate([f.reshape(11, -1) for f in L["fabs"]], axis=1)
    p = allv[po.NAMES.index("pressure")]
    post = np.isclose(allv[0], 8.0)
    assert post.any() and (~post).any()
    assert np.allclose(p[post], 116.5, rtol=1e-13) and np.allclose(p[~post], 1.0, rtol=1e-13)
    ux = allv[po.NAMES.index("ux")]
    assert np.allclose(ux[post], 8.25 * np.cos(np.pi / 6), rtol=1e-13) and np.allclose(ux[~post], 0.0)


def _lib_compare(a, b, **kw):
    from ocfd.plotfile import fcompare
    return fcompare(a, b, **kw)


def _report_errors(report):
    out = {}
    for line in report.splitlines():
        m = re.match(r"^ (\S+)\s+(\S+)\s+(\S+)\s*$", line)
        if m and m.group(1) in po.NAMES:
            out[m.group(1)] = (float(m.group(2)), float(m.group(3)))
    return out


def test_fcompare_identical(orc, tmp_path):
    boxes, A, og = _dmr_level(orc)
    po.write(tmp_path / "a", [(boxes, A, og)], 0.0, [0], 1.4)
    po.write(tmp_path / "b", [(boxes, A, og)], 0.0, [0], 1.4)
    res, rep = _lib_compare(tmp_path / "a", tmp_path / "b")
    assert res == 1 and "PLOTFILE AGREE" in rep and " level = 0" in rep
    assert all(v == (0.0, 0.0) for v in _report_errors(rep).values())


@pytest.mark.parametrize("norm", [0, 1, 2])
def test_fcompare_norms(orc, tmp_path, norm):
    boxes, A, og = _dmr_level(orc)
    po.write(tmp_path / "a", [(boxes, A, og)], 0.0, [0], 1.4)
    rng = np.random.default_rng(norm)
    for (lo, hi), f in zip(boxes, A):
        v = f.valid(lo, hi)
        v *= 1.0 + 1e-3 * rng.uniform(-1, 1, v.shape)
    po.write(tmp_path / "b", [(boxes, A, og)], 0.0, [0], 1.4)
    res, rep = _lib_compare(tmp_path / "a", tmp_path / "b", norm=norm)
    assert res == 0
    got = _report_errors(rep)
    want = po.compare(tmp_path / "a", tmp_path / "b", norm)[0]
    for name in ("rho", "xmom", "Eden", "ux", "pressure"):
        for g_, w in zip(got[name], want[name]):
            assert abs(g_ - w) <= 1e-9 * abs(w), (name, got[name], want[name])
    assert got["Cost"] == (0.0, 0.0)
    res2, rep2 = _lib_compare(tmp_path / "a", tmp_path / "b", norm=norm, rtol=1e-2)
    assert res2 == 2 and "PLOTFILE AGREE to specified tolerances" in rep2


def test_fcompare_nan_and_grids(orc, tmp_path):
    boxes, A, og = _dmr_level(orc)
    po.write(tmp_path / "a", [(boxes, A, og)], 0.0, [0], 1.4)
    boxes2, A2, _ = _dmr_level(orc, max_grid=(8, 8, 2))
    po.write(tmp_path / "c", [(boxes2, A2, og)], 0.0, [0], 1.4)
    with pytest.raises(RuntimeError, match="grids do not match"):
        _lib_compare(tmp_path / "a", tmp_path / "c")
    res, rep = _lib_compare(tmp_path / "a", tmp_path / "c", allow_diff_grids=True)
    assert res == 1
    A[0].valid(*boxes[0])[0, 0, 0, 0] = np.nan
    po.write(tmp_path / "d", [(boxes, A, og)], 0.0, [0], 1.4)
    res, rep = _lib_compare(tmp_path / "a", tmp_path / "d")
    assert res == -1 and "< NaN present in B >" in rep
