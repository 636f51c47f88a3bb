"""CPU restatement of the plotfile format and of fcompare -- test
infrastructure only (used by tests/ as the checker).

Writer: the AMReX "HyperCLaw-V1.1" plotfile of AmrLevel::writePlotFile as
restated in the reference's Source/IO.cpp:32-262 (Header, per-level VisMF
header Cell_H, data file Cell_D_00000 of FABs with one-line headers and raw
little-endian doubles), with the plot variables of EBR: the conserved state
(Source/Setup.cpp:202-206), Cost, the derived velocity and pressure
(ebr_dervel / ebr_derpres, Source/Setup.cpp:94-179) and vfrac.

Reader and comparison: fcompare, Util/Plotfile/fcompare.cpp:110-330 (norm 0
max, norm 1 and 2 scaled by dv^(1/norm), relative to plotfile A).
"""
import os

import numpy as np

NAMES = ("rho", "xmom", "ymom", "zmom", "Eden", "Cost", "ux", "uy", "uz", "pressure", "vfrac")
REAL_DESC = "(8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1))"


def _g(x):
    return "%.17g" % x


def _box(lo, hi):
    return "((%d,%d,%d) (%d,%d,%d) (0,0,0))" % (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])


def plot_vars(state, gamma, vfrac=1.0):
    """(11, nz, ny, nx) plot variables of a valid-region state (5, nz, ny, nx)."""
    r, mx, my, mz, E = state
    p = (gamma - 1.0) * (E - 0.5 / r * (mx * mx + my * my + mz * mz))
    return np.stack([r, mx, my, mz, E, np.zeros_like(r), mx / r, my / r, mz / r, p, np.full_like(r, vfrac)])


def write(path, levels, time, level_steps, gamma):
    """levels: list of (boxes, fabs, geom) from coarse to fine, fabs holding
    4 ghost cells (oracle Fab objects)."""
    os.makedirs(path, exist_ok=True)
    for l, (boxes, fabs, g) in enumerate(levels):
        ldir = os.path.join(path, "Level_%d" % l)
        os.makedirs(ldir, exist_ok=True)
        offs, mins, maxs = [], [], []
        with open(os.path.join(ldir, "Cell_D_00000"), "wb") as f:
            for (lo, hi), fab in zip(boxes, fabs):
                v = plot_vars(fab.valid(lo, hi), gamma)
                offs.append(f.tell())
                f.write(("FAB (%s)%s %d\n" % (REAL_DESC, _box(lo, hi), len(NAMES))).encode())
                f.write(np.ascontiguousarray(v, dtype="<f8").tobytes())
                mins.append(v.reshape(len(NAMES), -1).min(1))
                maxs.append(v.reshape(len(NAMES), -1).max(1))
        with open(os.path.join(ldir, "Cell_H"), "w") as f:
            f.write("1\n1\n%d\n0\n" % len(NAMES))
            f.write("(%d 0\n" % len(boxes))
            for lo, hi in boxes:
                f.write(_box(lo, hi) + "\n")
            f.write(")\n%d\n" % len(boxes))
            for o in offs:
                f.write("FabOnDisk: Cell_D_00000 %d\n" % o)
            f.write("\n")
            for arr in (mins, maxs):
                f.write("%d,%d\n" % (len(boxes), len(NAMES)))
                for row in arr:
                    f.write("".join(_g(x) + "," for x in row) + "\n")
                f.write("\n")
    g0 = levels[0][2]
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("HyperCLaw-V1.1\n%d\n" % len(NAMES))
        for n in NAMES:
            f.write(n + "\n")
        f.write("3\n%s\n%d\n" % (_g(time), len(levels) - 1))
        f.write("".join(_g(g0.prob_lo[d]) + " " for d in range(3)) + "\n")
        f.write("".join(_g(g0.prob_lo[d] + (g0.dom_hi[d] - g0.dom_lo[d] + 1) * g0.dx[d]) + " " for d in range(3)) + "\n")
        f.write("".join("2 " for _ in range(len(levels) - 1)) + "\n")
        f.write("".join(_box(list(g.dom_lo), list(g.dom_hi)) + " " for _, _, g in levels) + "\n")
        f.write("".join("%d " % s for s in level_steps) + "\n")
        for _, _, g in levels:
            f.write("".join(_g(g.dx[d]) + " " for d in range(3)) + "\n")
        f.write("0\n0\n")
        for l, (boxes, _, g) in enumerate(levels):
            f.write("%d %d %s\n%d\n" % (l, len(boxes), _g(time), level_steps[l]))
            for lo, hi in boxes:
                for d in range(3):
                    f.write("%s %s\n" % (_g(g.prob_lo[d] + lo[d] * g.dx[d]), _g(g.prob_lo[d] + (hi[d] + 1) * g.dx[d])))
            f.write("Level_%d/Cell\n" % l)


def read(path):
    """{'names', 'time', 'levels': [{'dx', 'dom': (lo, hi), 'boxes', 'fabs'}]}"""
    with open(os.path.join(path, "Header")) as f:
        tok = f.read().split("\n")
    assert tok[0] == "HyperCLaw-V1.1"
    nv = int(tok[1])
    names = tok[2:2 + nv]
    rest = " ".join(tok[2 + nv:]).split()
    dim, time, finest = int(rest[0]), float(rest[1]), int(rest[2])
    assert dim == 3
    q = 3 + 6 + finest
    doms = []
    for _ in range(finest + 1):
        s = rest[q] + rest[q + 1] + rest[q + 2]
        nums = [int(x) for x in s.replace("(", " ").replace(")", " ").replace(",", " ").split()]
        doms.append((nums[0:3], nums[3:6]))
        q += 3
    q += finest + 1
    dxs = []
    for _ in range(finest + 1):
        dxs.append([float(x) for x in rest[q:q + 3]])
        q += 3
    levels = []
    for l in range(finest + 1):
        ldir = os.path.join(path, "Level_%d" % l)
        with open(os.path.join(ldir, "Cell_H")) as f:
            lines = f.read().split("\n")
        nb = int(lines[4][1:].split()[0])
        boxes = []
        for b in range(nb):
            nums = [int(x) for x in lines[5 + b].replace("(", " ").replace(")", " ").replace(",", " ").split()]
            boxes.append((nums[0:3], nums[3:6]))
        fod = [ln.split() for ln in lines[7 + nb:7 + 2 * nb]]
        fabs = []
        for (lo, hi), (_, fname, off) in zip(boxes, fod):
            with open(os.path.join(ldir, fname), "rb") as f:
                f.seek(int(off))
                hdr = f.readline().decode()
                assert hdr.startswith("FAB (") and REAL_DESC in hdr
                shape = (nv,) + tuple(hi[d] - lo[d] + 1 for d in (2, 1, 0))
                fabs.append(np.frombuffer(f.read(8 * int(np.prod(shape))), dtype="<f8").reshape(shape))
        levels.append({"dx": dxs[l], "dom": doms[l], "boxes": boxes, "fabs": fabs})
    return {"names": names, "time": time, "levels": levels}


def compare(a, b, norm=0):
    """Per level, per variable of A: (absolute, relative) errors as fcompare
    computes them, for plotfiles on the same grids."""
    A, B = read(a), read(b)
    out = []
    for La, Lb in zip(A["levels"], B["levels"]):
        dv = float(np.prod(La["dx"]))
        res = {}
        for ia, name in enumerate(A["names"]):
            ib = B["names"].index(name)
            da = np.concatenate([f[ia].ravel() for f in La["fabs"]])
            db = np.concatenate([f[ib].ravel() for f in Lb["fabs"]])
            e = np.abs(db - da)
            with np.errstate(invalid="ignore", divide="ignore"):
                if norm == 0:
                    res[name] = (e.max(), e.max() / np.abs(da).max())
                elif norm == 1:
                    res[name] = (e.sum() * dv, e.sum() / np.abs(da).sum())
                else:
                    res[name] = (np.sqrt((e * e).sum()) * np.sqrt(dv),
                                 np.sqrt((e * e).sum()) / np.sqrt((da * da).sum()))
        out.append(res)
    return out
