"""Shared helpers of the GPU parity tests: move level data between the HIP
library's MultiFab storage and the oracle's per-box FABs."""
import ctypes

import numpy as np


def to_ogeom(g):
    import oracle
    o = oracle.OGeom()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(g), ctypes.sizeof(o))
    return o


def upload(level, mf, fabs, ngrow=4):
    import torch
    for b, f in enumerate(fabs):
        level.fab_view(mf, b, ngrow).copy_(torch.from_numpy(f.a))


def download(level, mf, ngrow=4):
    return [level.fab_view(mf, b, ngrow).cpu().numpy().copy() for b in range(level.nbox)]


def random_state(fab, rng, kind="smooth", vmax=0.5, gamma=1.4):
    """Fill a FAB (grown box) with a positive-density, positive-pressure state."""
    nc, nz, ny, nx = fab.a.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if kind == "smooth":
        ph = rng.uniform(0, 2 * np.pi, size=8)
        rho = 1.0 + 0.3 * np.sin(0.3 * x + ph[0]) * np.cos(0.2 * y + ph[1]) + 0.1 * np.sin(0.25 * z + ph[2])
        p = 1.0 + 0.2 * np.cos(0.17 * x + ph[3]) + 0.1 * np.sin(0.31 * y + 0.2 * z + ph[4])
        u = vmax * np.sin(0.21 * x + 0.1 * y + ph[5])
        v = vmax * np.cos(0.13 * y + 0.2 * z + ph[6])
        w = vmax * np.sin(0.19 * z + 0.3 * x + ph[7])
    else:
        rho = rng.uniform(0.1, 2.0, size=x.shape)
        p = rng.uniform(0.1, 2.0, size=x.shape)
        u = rng.uniform(-vmax, vmax, size=x.shape) * 3
        v = rng.uniform(-vmax, vmax, size=x.shape) * 3
        w = rng.uniform(-vmax, vmax, size=x.shape) * 3
    fab.a[0] = rho
    fab.a[1] = rho * u
    fab.a[2] = rho * v
    fab.a[3] = rho * w
    fab.a[4] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)


def rel_err(a, b):
    """(L1, Linf) of a-b relative to b, per component (first axis)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for n in range(a.shape[0]):
        d = np.abs(a[n] - b[n])
        den1 = max(np.abs(b[n]).sum(), 1e-300)
        deninf = max(np.abs(b[n]).max(), 1e-300)
        out.append((d.sum() / den1, d.max() / deninf))
    return out


def max_rel(a, b):
    return level_max_rel([a], [b])


def level_rel_err(got, ref):
    """Whole-level (L1, Linf) error per conserved variable: sum/max over all
    boxes of |got-ref| over sum/max of |ref|.  The momentum components share
    the norm of the momentum vector, so a component that is identically zero
    (e.g. z-momentum of a planar case) is measured against the flow's momentum."""
    nc = got[0].shape[0]
    num1 = np.zeros(nc); numi = np.zeros(nc); den1 = np.zeros(nc); deni = np.zeros(nc)
    for a, b in zip(got, ref):
        a = np.asarray(a).reshape(nc, -1)
        b = np.asarray(b).reshape(nc, -1)
        d = np.abs(a - b)
        num1 += d.sum(1); numi = np.maximum(numi, d.max(1))
        den1 += np.abs(b).sum(1); deni = np.maximum(deni, np.abs(b).max(1))
    if nc >= 4:
        den1[1:4] = den1[1:4].sum()
        deni[1:4] = deni[1:4].max()
    den1 = np.maximum(den1, 1e-300); deni = np.maximum(deni, 1e-300)
    return list(zip(num1 / den1, numi / deni))


def level_max_rel(got, ref):
    return max(max(e) for e in level_rel_err(got, ref))
