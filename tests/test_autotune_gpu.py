"""The march tile-shape autotuner: it times every candidate chunk length on
the level and keeps the fastest, and the chunk length never changes results
(every candidate gives the same state bit for bit)."""
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu


def test_chunk_length_does_not_change_results(torch_cuda):
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.kh3d(32, max_grid=(32, 32, 16))
    p = ocfd.default_parm()
    flags = ocfd.VISC | ocfd.FLUX_FP32
    out = []
    for ch in (8, 16, 32, 64):
        lev = ocfd.Level(boxes, g)
        assert lev.march_chunk(ch) == ch
        A, B, C = lev.new_mf(), lev.new_mf(), lev.new_mf()
        lev.init_state(kind, A, p, seed=4)
        ocfd.run(lev, A, B, C, p, 0.5, 2, flags, nsteps=3)
        out.append(gu.download(lev, A))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


def test_autotune_picks_a_timed_candidate(torch_cuda):
    import ocfd
    from ocfd import cases
    boxes, g, kind = cases.kh3d(64)
    lev = ocfd.Level(boxes, g)
    p = ocfd.default_parm()
    A = lev.new_mf()
    lev.init_state(kind, A, p, seed=4)
    lev.fill_boundary(A, 0.0)
    ch, ms = lev.autotune(A, p, ocfd.VISC | ocfd.FLUX_FP32, reps=3)
    print("march chunk", ch, "timings (ms)", ms)
    assert ch in ms and all(v > 0 for v in ms.values())
    assert ms[ch] == min(ms.values())
    assert lev.march_chunk() == ch
