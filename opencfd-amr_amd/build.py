"""Build the HIP library of the per-box advance (libocfd.so, in-tree).

Every HIP source is compiled for gfx950 with hipcc; packed FP32 instructions
are disabled for the device code.  The library is linked in place next to this
file so it ships with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libocfd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


# per-file extra flags: the chemistry follows the reference's FP64 operation
# order without FMA contraction
EXTRA = {"ocfd_chem.hip": ["-ffp-contract=off"]}


def sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(str(c) for c in cmd))
    return r.stdout


def _jobs():
    # at most 16 compilers at once, fewer where the machine has fewer cores
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n, int(os.environ.get("MAX_JOBS", "16"))))


def build(verbose=False):
    BUILD.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in sources():
        obj = BUILD / (src.name + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        cmds.append([HIPCC] + HIP_FLAGS + EXTRA.get(src.name, []) + lang + ["-c", str(src), "-o", str(obj)])
        objs.append(str(obj))
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        list(ex.map(_run, cmds))
    tmp = str(LIB) + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


DRIVER = PKG / "ocfd_run"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))


def driver_cmd(src, out):
    """The host-only compile of a C++ program over include/ocfd.hpp, linked
    against libocfd.so (found next to the program at run time)."""
    return [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
            f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}", str(src), f"-L{PKG}", "-locfd",
            f"-L{ROCM / 'lib'}", "-lamdhip64", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}", "-o", str(out)]


def build_driver(verbose=False):
    """ocfd_run: the C++ host driver of one level (host/ocfd_run.cpp)."""
    cmd = driver_cmd(PKG / "host" / "ocfd_run.cpp", str(DRIVER) + ".tmp")
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(str(DRIVER) + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
