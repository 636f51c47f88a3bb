"""Box lists and geometries of the configured workloads (the Exec/<case>/inputs
of the reference, restated as level set-ups)."""
import numpy as np

from . import (Geom, Level, make_geom, dmr_states, eb_cylinder, inflow_state, INTERIOR, INFLOW, OUTFLOW, SLIPWALL,
               CASE_SOD, CASE_DMR, CASE_KH3D, CASE_EB_INFLOW, CASE_CHEMSOD)


def chop(n_cell, max_grid):
    """Box list covering [0, n_cell) with boxes of at most max_grid per side
    (amr.max_grid_size); the domain is split into equal-as-possible pieces."""
    edges = []
    for d in range(3):
        n = n_cell[d]
        m = max(1, -(-n // max_grid[d]))
        cuts = [n * q // m for q in range(m + 1)]
        edges.append(cuts)
    boxes = []
    for k in range(len(edges[2]) - 1):
        for j in range(len(edges[1]) - 1):
            for i in range(len(edges[0]) - 1):
                lo = (edges[0][i], edges[1][j], edges[2][k])
                hi = (edges[0][i + 1] - 1, edges[1][j + 1] - 1, edges[2][k + 1] - 1)
                boxes.append((lo, hi))
    return boxes


def uneven_chop(n_cell, sizes, seed):
    """Box list covering [0, n_cell) in x and y with seeded uneven box sizes
    drawn from `sizes` (e.g. max_grid_size 32..128); z is not split."""
    rng = np.random.default_rng(seed)
    edges = []
    for d in range(2):
        cuts = [0]
        while cuts[-1] < n_cell[d]:
            s = int(rng.choice(sizes))
            cuts.append(min(n_cell[d], cuts[-1] + s))
        edges.append(cuts)
    boxes = []
    for j in range(len(edges[1]) - 1):
        for i in range(len(edges[0]) - 1):
            boxes.append(((edges[0][i], edges[1][j], 0),
                          (edges[0][i + 1] - 1, edges[1][j + 1] - 1, n_cell[2] - 1)))
    return boxes


def sod(nx=512, ny=8, nz=8, max_grid=None):
    """2D Sod shock tube as a z-uniform slab: [0,1] x [0,ny/nx] x [0,nz/nx],
    x outflow, y/z slip walls (Exec/0_SOD/inputs:4-16)."""
    n = (nx, ny, nz)
    g = make_geom(n, (0.0, 0.0, 0.0), (1.0, ny / nx, nz / nx), periodic=(0, 0, 0),
                  lo_bc=(OUTFLOW, SLIPWALL, SLIPWALL), hi_bc=(OUTFLOW, SLIPWALL, SLIPWALL))
    boxes = chop(n, max_grid or n)
    return boxes, g, CASE_SOD


def dmr(nx=1024, ny=256, nz=2, sizes=None, seed=1234, max_grid=None):
    """Double Mach reflection on [0,4] x [0,1] x [0,nz/256]: inflow x-lo,
    slip wall y-lo, outflow x-hi, time-dependent inflow y-hi
    (Exec/1_DMR/inputs:4-16, prob_bcfill.cpp)."""
    n = (nx, ny, nz)
    left, right = dmr_states()
    g = make_geom(n, (0.0, 0.0, 0.0), (4.0, 4.0 * ny / nx, 4.0 * nz / nx), periodic=(0, 0, 0),
                  lo_bc=(INFLOW, SLIPWALL, OUTFLOW), hi_bc=(OUTFLOW, INFLOW, OUTFLOW), probtype=1,
                  left_state=left, right_state=right)
    if sizes is not None:
        boxes = uneven_chop(n, sizes, seed)
    else:
        boxes = chop(n, max_grid or n)
    return boxes, g, CASE_DMR


def kh3d(n=256, max_grid=None):
    """3D viscous Kelvin-Helmholtz shear layer, periodic unit cube."""
    nc = (n, n, n)
    g = make_geom(nc, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), periodic=(1, 1, 1),
                  lo_bc=(INTERIOR,) * 3, hi_bc=(INTERIOR,) * 3)
    boxes = chop(nc, max_grid or nc)
    return boxes, g, CASE_KH3D


def eb_channel(nx=512, ny=128, nz=2, max_grid=32, radius=0.1, center=(0.5, 0.25), small_volfrac=0.05):
    """Supersonic channel past a cylinder (diameter 0.2 in a 2 x 0.5 channel):
    Mach 3 inflow at x-lo (the inflow state of Exec/3_FSTEP), outflow at x-hi,
    slip walls in y, periodic in z.  Cells under small_volfrac are covered
    (eb2.small_volfrac of Exec/Corner/inputs); run it with OCFD_EB_REDIST, the
    small-cell scaling that keeps the explicit cut-cell update stable.
    Returns (boxes, geom, kind, eb spec)."""
    n = (nx, ny, nz)
    dx = 2.0 / nx
    st = inflow_state()
    g = make_geom(n, (0.0, 0.0, 0.0), (2.0, 0.5 * ny / (nx / 4), nz * dx), periodic=(0, 0, 1),
                  lo_bc=(INFLOW, SLIPWALL, INTERIOR), hi_bc=(OUTFLOW, SLIPWALL, INTERIOR), probtype=2,
                  left_state=st, right_state=st)
    boxes = chop(n, (max_grid, max_grid, nz))
    return boxes, g, CASE_EB_INFLOW, eb_cylinder(center[0], center[1], radius, small_volfrac)


def chem_channel(nx=512, ny=128, nz=2, max_grid=32, T0=1100.0, p0=101325.0, phi=1.0, u0=200.0, radius=0.1,
                 center=(0.5, 0.25), small_volfrac=0.05):
    """Reacting premixed H2/air channel past a cylinder: a 2 x 0.5 m channel
    (nx x ny cells, nz periodic z layers), a cylinder of diameter 0.2 m,
    premixed stoichiometric H2/air at 1100 K and 1 atm flowing in at x-lo at
    200 m/s (constant inflow of the same mixture), outflow at x-hi, slip walls
    in y.  9 species, the multispecies path with the EB-CHEM cut boxes.
    The initial data fill the whole channel with the inflow mixture.
    Returns (boxes, geom, kind, eb spec, (state, partial densities))."""
    from . import premixed_state, CASE_CHEM_CHANNEL
    n = (nx, ny, nz)
    dx = 2.0 / nx
    state, spec = premixed_state(T0, p0, phi, u0)
    g = make_geom(n, (0.0, 0.0, 0.0), (2.0, 0.5 * ny / (nx / 4), nz * dx), periodic=(0, 0, 1),
                  lo_bc=(INFLOW, SLIPWALL, INTERIOR), hi_bc=(OUTFLOW, SLIPWALL, INTERIOR), probtype=2,
                  left_state=state, right_state=state)
    boxes = chop(n, (max_grid, max_grid, nz))
    return boxes, g, CASE_CHEM_CHANNEL, eb_cylinder(center[0], center[1], radius, small_volfrac), (state, spec)


def chemsod(nx=1280, ny=4, nz=4, max_grid=128):
    """H2/O2 shock tube with a hot high-pressure driver (Exec/5_CHEMSOD/inputs):
    dx = 1e-4 m, slip wall at x-lo, outflow at x-hi, periodic in y and z."""
    n = (nx, ny, nz)
    dx = 1e-4
    g = make_geom(n, (0.0, 0.0, 0.0), (nx * dx, ny * dx, nz * dx), periodic=(0, 1, 1),
                  lo_bc=(SLIPWALL, INTERIOR, INTERIOR), hi_bc=(OUTFLOW, INTERIOR, INTERIOR))
    return chop(n, (max_grid, max_grid, max_grid)), g, CASE_CHEMSOD
