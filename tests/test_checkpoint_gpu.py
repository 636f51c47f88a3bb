"""Checkpoint and restart of an AMR hierarchy (Amr::checkPoint /
Amr::restart; the per-level state as AmrLevel::checkPoint writes it,
Source/IO.cpp:9-30).

A run interrupted by a checkpoint and continued from a restart must equal the
uninterrupted run bit for bit: the grids of every level, the clock (time,
dt per level, level steps and regrid counters) and the states.  The level
files are VisMF MultiFabs whose FABs hold exactly the valid cells of the
state, read back here with a plain numpy reader of the format."""
import os

import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu


def _dmr_hier(**over):
    import ocfd
    from ocfd import amr, cases
    boxes, g, kind = cases.dmr(64, 16, 2, max_grid=(16, 16, 2))
    kw = dict(scheme=2, max_level=2, refine_dengrad=0.05, regrid_int=2, blocking_factor=2, max_grid_size=32)
    kw.update(over)
    return boxes, g, kw, amr.Hierarchy(boxes, g, ocfd.default_parm(), 0.5, kind, **kw)


def _read_vismf(base):
    with open(base + "_H") as f:
        lines = f.read().split("\n")
    assert lines[:4] == ["1", "1", "5", "0"]
    nb = int(lines[4][1:].split()[0])
    boxes = []
    for b in range(nb):
        n = [int(x) for x in lines[5 + b].replace("(", " ").replace(")", " ").replace(",", " ").split()]
        boxes.append((tuple(n[0:3]), tuple(n[3:6])))
    fabs = []
    for (lo, hi), ln in zip(boxes, lines[7 + nb:7 + 2 * nb]):
        _, fname, off = ln.split()
        with open(os.path.join(os.path.dirname(base), fname), "rb") as f:
            f.seek(int(off))
            assert f.readline().startswith(b"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
            shape = (5,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
            fabs.append(np.frombuffer(f.read(8 * int(np.prod(shape))), dtype="<f8").reshape(shape))
    return boxes, fabs


def test_checkpoint_files_hold_the_valid_state(torch_cuda, tmp_path):
    _, _, _, H = _dmr_hier()
    H.run(2)
    H.checkpoint(str(tmp_path / "chk"))
    assert H.finest >= 1
    for l in range(H.finest + 1):
        boxes, fabs = _read_vismf(str(tmp_path / "chk" / ("Level_%d" % l) / "SD_0_New"))
        assert boxes == [tuple(map(tuple, b)) for b in H.levels[l].boxes]
        for a, got in zip(gu.download(H.levels[l], H.S[l]), fabs):
            assert np.array_equal(a[:, 4:-4, 4:-4, 4:-4], got)


@pytest.mark.parametrize("split", [1, 2, 3])
def test_restart_continues_bitwise(torch_cuda, tmp_path, split):
    """Three-level double Mach reflection, regrid every 2 steps: 5 coarse steps
    in one go against `split` steps, checkpoint, restart, 5 - split steps."""
    import ocfd
    from ocfd import amr
    boxes, g, kw, A = _dmr_hier()
    A.run(5)
    _, _, _, B = _dmr_hier()
    B.run(split)
    B.checkpoint(str(tmp_path / "chk"))
    del B
    C = amr.Hierarchy.restart(str(tmp_path / "chk"), g, ocfd.default_parm(), 0.5,
                              **{k: v for k, v in kw.items() if k != "max_level"})
    C.run(5 - split)
    assert (C.time, C.steps, C.finest) == (A.time, A.steps, A.finest)
    assert C.dt_level == A.dt_level and C.level_steps == A.level_steps and C.count == A.count
    for l in range(A.finest + 1):
        assert C.levels[l].boxes == A.levels[l].boxes, l
        for a, c in zip(gu.download(A.levels[l], A.S[l]), gu.download(C.levels[l], C.S[l])):
            assert np.array_equal(a[:, 4:-4, 4:-4, 4:-4], c[:, 4:-4, 4:-4, 4:-4]), l


def test_restart_rejects_other_grids(torch_cuda, tmp_path):
    import ocfd
    from ocfd import amr
    _, g, kw, H = _dmr_hier()
    H.checkpoint(str(tmp_path / "chk"))
    hdr = (tmp_path / "chk" / "Header").read_text().split("\n")
    i = hdr.index("level 0 %d" % len(H.levels[0].boxes))
    lo_hi = hdr[i + 1].split()
    lo_hi[3] = str(int(lo_hi[3]) - 1)          # shrink the first box: no longer the saved box
    hdr[i + 1] = " ".join(lo_hi)
    (tmp_path / "chk" / "Header").write_text("\n".join(hdr))
    with pytest.raises(RuntimeError, match="differs from the level"):
        amr.Hierarchy.restart(str(tmp_path / "chk"), g, ocfd.default_parm(), 0.5, **kw)
    with pytest.raises(ValueError, match="below the checkpoint"):
        amr.Hierarchy.restart(str(tmp_path / "chk"), g, ocfd.default_parm(), 0.5, **dict(kw, max_level=0))
