"""Plotfiles: the AMReX "HyperCLaw-V1.1" format the reference writes
(Source/EBR.H:33-35, Source/IO.cpp:32-262) and the fcompare comparison
(Util/Plotfile/fcompare.cpp), both in the HIP library."""
import ctypes

from . import _stream, check, lib

PLOT_VARS = ("rho", "xmom", "ymom", "zmom", "Eden", "Cost", "ux", "uy", "uz", "pressure", "vfrac")


def write_plotfile(path, levels, states, time, level_steps, parm, stream=None):
    """levels: coarse to fine (each refines the one below by 2); states: the
    matching 4-ghost state MultiFabs (device tensors)."""
    n = len(levels)
    if n == 0 or len(states) != n or len(level_steps) != n:
        raise ValueError("write_plotfile needs one state and one step count per level")
    hs = (ctypes.c_void_p * n)(*[lv.handle for lv in levels])
    ss = (ctypes.c_void_p * n)(*[s.data_ptr() for s in states])
    st = (ctypes.c_int * n)(*level_steps)
    check(lib().ocfd_plotfile_write(str(path).encode(), n, hs, ss, time, st, ctypes.byref(parm), _stream(stream)),
          "ocfd_plotfile_write")


def fcompare(a, b, norm=0, rtol=0.0, atol=0.0, allow_diff_grids=False):
    """(result, report): result 1 identical, 2 within tolerances, 0 different,
    -1 NaN present; report is fcompare's table."""
    res = ctypes.c_int()
    buf = ctypes.create_string_buffer(1 << 20)
    check(lib().ocfd_plotfile_compare(str(a).encode(), str(b).encode(), norm, rtol, atol, int(allow_diff_grids),
                                      ctypes.byref(res), buf, len(buf)), "ocfd_plotfile_compare")
    return res.value, buf.value.decode()
