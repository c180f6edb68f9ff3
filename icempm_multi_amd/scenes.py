"""Seeded synthetic ice scenes in the reference's input layout.

The reference reads point clouds (x, y, uint64 llGrainIDs, float32 `volume`
attribute) from HDF5 and prepares them in SnapshotManager::LoadRawPoints
(simulation/snapshotmanager.cpp:29-121).  These generators produce the same
four inputs from a SplitMix64 stream and hand them to the library's
``prepare_raw_points`` (the LoadRawPoints logic after the file read).

Workloads (all FP64 inputs; sizes in cells of side h):
  block   40 x 10 m block, grid 256 x 128, h = 0.25, 4 jittered points/cell
  sheet   200 x 12.5 m sheet, grid 4096 x 1024, h = 0.05, Poisson-disk cloud of
          4 points/cell on average (or jittered), cylindrical indenter D = 1 m,
          depth 0.25 m, moving at 2 m/s
  batch   many blocks of log-uniform point counts and per-scene E, cohesion
  floes   fields of seeded convex polygonal floes, power-law sizes
"""
import math

import numpy as np

from . import capi
from .params import N_PTS_ARRAYS, SimParams

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) after state `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """n doubles in [0, 1) from SplitMix64 (53-bit mantissa)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def make_params(**kw):
    """SimParams::Reset, then the given keys (JSON names), then the derived constants."""
    P = capi.params_reset()
    alias = {"GridX": "GridXTotal"}
    for k, v in kw.items():
        if k == "DP_phi":
            P.DP_tan_phi = math.tan(v * math.pi / 180)
            continue
        setattr(P, alias.get(k, k), v)
    return capi.params_derive(P)


def jittered_block(nx_cells, ny_cells, h, seed, per_side=2, jitter=0.25, x0=0.0, y0=0.0):
    """per_side^2 points per cell on a sub-lattice, each moved by U(-jitter, jitter) sub-spacings."""
    sub = h / per_side
    i = np.arange(nx_cells * per_side, dtype=np.float64)
    j = np.arange(ny_cells * per_side, dtype=np.float64)
    X, Y = np.meshgrid(i, j, indexing="ij")
    n = X.size
    u = uniform(seed, 2 * n)
    x = x0 + (X.ravel() + 0.5 + (u[:n] - 0.5) * 2 * jitter) * sub
    y = y0 + (Y.ravel() + 0.5 + (u[n:] - 0.5) * 2 * jitter) * sub
    return x, y


# Points per r^2 of area that the Poisson-disk sampler fills (measured on
# large boxes; the reference's generator uses 0.656 with its sampler).
POISSON_FILL = 0.62


def poisson_block(nx_cells, ny_cells, h, seed, per_cell=4, x0=0.0, y0=0.0):
    """Poisson-disk cloud of about per_cell points per cell over the block, as the
    reference's generator makes it (grainprocessor2d.cpp:40-56): minimum distance
    r = sqrt(POISSON_FILL * area / n), n = per_cell * cells (r = 0.39 h at 4 per cell)."""
    w, hgt = nx_cells * h, ny_cells * h
    r = math.sqrt(POISSON_FILL * w * hgt / (per_cell * nx_cells * ny_cells))
    return capi.poisson_disk(r, x0, y0, x0 + w, y0 + hgt, seed)


def voronoi_grains(x, y, n_grains, seed):
    """Grain id of each point: index of the nearest of n_grains seeded sites."""
    from scipy.spatial import cKDTree
    u = uniform(seed ^ 0x5EED, 2 * n_grains)
    sx = x.min() + u[:n_grains] * (x.max() - x.min())
    sy = y.min() + u[n_grains:] * (y.max() - y.min())
    _, idx = cKDTree(np.stack([sx, sy], 1)).query(np.stack([x, y], 1), workers=-1)
    return idx.astype(np.uint64)


def load_raw_points(P, x, y, grains, volume, extra_space=None):
    """LoadRawPoints (snapshotmanager.cpp:29-121) on in-memory arrays: HostSideSOA
    of capacity n*(1+ExtraSpaceForIncomingPoints), x/y/llGrainIDs in lines 4/5/0,
    then offset, sort, F = I, Jp_inv = 1, indenter start and particle volume."""
    n = len(x)
    if extra_space is None:
        extra_space = P.ExtraSpaceForIncomingPoints
    cap = int(n * (1 + extra_space))
    hssoa = np.zeros((N_PTS_ARRAYS, cap), dtype=np.float64)
    hssoa[4, :n] = x
    hssoa[5, :n] = y
    hssoa[0, :n] = np.asarray(grains, dtype=np.uint64).view(np.float64)
    P.nPtsTotal = n
    kept = capi.prepare_raw_points(P, hssoa, n, np.float32(volume))
    return hssoa, kept


def block_scene(seed=0x1CE0000, E=9e9, cohesion=1e5, phi=30.0, nx=160, ny=40, h=0.25, grid=(256, 128),
                dt=1e-5, halo=8, n_grains=1000, **extra):
    """The 40 x 10 m block (h = 0.25 m, 256 x 128 grid, 4 jittered points per cell, ~25.6k points)."""
    tan_phi = math.tan(phi * math.pi / 180)
    kw = dict(GridX=grid[0], GridY=grid[1], GridXDimension=grid[0] * h, InitialTimeStep=dt, YoungsModulus=E,
              PoissonsRatio=0.3, Density=980.0, DP_phi=phi, DP_threshold_p=cohesion / tan_phi, SetupType=2,
              IndVelocity=0.0, GridHaloSize=halo, nPartitions=1)
    kw.update(extra)
    P = make_params(**kw)
    x, y = jittered_block(nx, ny, h, seed)
    g = voronoi_grains(x, y, n_grains, seed)
    hssoa, n = load_raw_points(P, x, y, g, nx * h * ny * h)
    return P, hssoa, n


def sheet_scene(seed=0x1CE0001, E=9e9, cohesion=1e5, phi=30.0, nx=4000, ny=250, h=0.05, grid=(4096, 1024),
                dt=1e-5, halo=8, ind_diameter=1.0, ind_depth=0.25, ind_velocity=2.0, n_grains=1000,
                sampler="poisson", **extra):
    """The 200 x 12.5 m sheet (grid 4096 x 1024, h = 0.05 m, ~4.0M points) with a moving cylindrical
    indenter.  sampler: "poisson" (Poisson-disk, 4 points per cell on average) or "jittered"."""
    tan_phi = math.tan(phi * math.pi / 180)
    kw = dict(GridX=grid[0], GridY=grid[1], GridXDimension=grid[0] * h, InitialTimeStep=dt, YoungsModulus=E,
              PoissonsRatio=0.3, Density=980.0, DP_phi=phi, DP_threshold_p=cohesion / tan_phi, SetupType=0,
              IndDiameter=ind_diameter, IndDepth=ind_depth, IndVelocity=ind_velocity, GridHaloSize=halo,
              nPartitions=1)
    kw.update(extra)
    P = make_params(**kw)
    if sampler == "poisson":
        x, y = poisson_block(nx, ny, h, seed)
    elif sampler == "jittered":
        x, y = jittered_block(nx, ny, h, seed)
    else:
        raise ValueError(f"unknown sampler {sampler!r}")
    g = voronoi_grains(x, y, n_grains, seed)
    hssoa, n = load_raw_points(P, x, y, g, nx * h * ny * h)
    return P, hssoa, n


_GRID_X = (128, 256, 512, 1024)
_GRID_Y = (128, 256, 512)


def batch_scenes(n_scenes=64, seed=0x1CE0100, n_min=10_000, n_max=500_000, h=0.1, dt=1e-5, halo=8):
    """Independent blocks: point counts log-uniform in [n_min, n_max], grids from
    128x128 to 1024x512 chosen to fit, E ~ U[3e9, 9e9], cohesion ~ U[5e4, 2e5]."""
    out = []
    for s in range(n_scenes):
        u = uniform(seed + s, 4)
        n_target = int(round(math.exp(math.log(n_min) + u[0] * (math.log(n_max) - math.log(n_min)))))
        E = 3e9 + u[1] * 6e9
        coh = 5e4 + u[2] * 1.5e5
        cells = n_target / 4.0
        by = max(4, int(round(math.sqrt(cells / 4.0))))
        bx = max(4, int(round(cells / by)))
        gx = next((g for g in _GRID_X if g >= bx * 1.1 + 2 * halo + 8), _GRID_X[-1])
        gy = next((g for g in _GRID_Y if g >= by + 16), _GRID_Y[-1])
        bx = min(bx, gx - 2 * halo - 8)
        by = min(by, gy - 16)
        P, hssoa, n = block_scene(seed=seed + 1000 + s, E=E, cohesion=coh, nx=bx, ny=by, h=h, grid=(gx, gy),
                                  dt=dt, halo=halo, n_grains=200)
        out.append((P, hssoa, n))
    return out


def _convex_polygon(cx, cy, r, u):
    """Convex polygon around (cx, cy): sorted random angles on a circle of radius r."""
    k = 5 + int(u[0] * 6)
    ang = np.sort(u[1:1 + k] * 2 * math.pi)
    return np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1)


def _inside_convex(px, py, poly):
    inside = np.ones(px.shape, dtype=bool)
    k = len(poly)
    for i in range(k):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % k]
        inside &= (x1 - x0) * (py - y0) - (y1 - y0) * (px - x0) >= 0
    return inside


def floe_field(seed=0x1CE0200, grid=(8192, 2048), h=0.05, target_points=3_000_000, r_min=0.5, r_max=10.0,
               dt=1e-5, halo=8, E=9e9, cohesion=1e5, phi=30.0):
    """A field of seeded convex polygonal floes (size ~ power law, exponent -2,
    1-20 m across) on a jittered 4-points-per-cell lattice."""
    tan_phi = math.tan(phi * math.pi / 180)
    P = make_params(GridX=grid[0], GridY=grid[1], GridXDimension=grid[0] * h, InitialTimeStep=dt,
                    YoungsModulus=E, PoissonsRatio=0.3, Density=980.0, DP_phi=phi,
                    DP_threshold_p=cohesion / tan_phi, SetupType=2, IndVelocity=0.0, GridHaloSize=halo,
                    nPartitions=1)
    W = (grid[0] - 2 * halo - 16) * h
    H = (grid[1] - 24) * h
    r_max = min(r_max, 0.4 * min(W, H))       # small grids: every floe fits inside
    r_min = min(r_min, 0.5 * r_max)
    area_target = target_points * (h * h) / 4.0
    xs, ys, gs = [], [], []
    area = 0.0
    k = 0
    occupied = []
    while area < area_target and k < 200000:
        u = uniform(seed * 7919 + k, 16)
        k += 1
        # power law p(r) ~ r^-2 on [r_min, r_max]
        r = 1.0 / (1.0 / r_min - u[0] * (1.0 / r_min - 1.0 / r_max))
        cx = r + u[1] * (W - 2 * r)
        cy = r + u[2] * (H - 2 * r)
        if any((cx - ox) ** 2 + (cy - oy) ** 2 < (r + orr + 2 * h) ** 2 for ox, oy, orr in occupied):
            continue
        occupied.append((cx, cy, r))
        poly = _convex_polygon(cx, cy, r, u[3:])
        n_side = int(math.ceil(2 * r / h))
        x, y = jittered_block(n_side, n_side, h, seed + 31 * k, x0=cx - r, y0=cy - r)
        m = _inside_convex(x, y, poly)
        xs.append(x[m])
        ys.append(y[m])
        gs.append(np.full(m.sum(), len(occupied) % 1000, dtype=np.uint64))
        area += m.sum() * h * h / 4.0
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    g = np.concatenate(gs)
    hssoa, n = load_raw_points(P, x, y, g, area)
    return P, hssoa, n


FLOE_GRIDS = ((8192, 2048), (8192, 1024), (4096, 2048), (4096, 1024))


def floe_fields(n_scenes=8, seed=0x1CE0500, total_points=24_000_000, grids=FLOE_GRIDS, h=0.05, dt=1e-5):
    """Independent floe fields (the combined workload): grids cycling through
    `grids` (up to 8192 x 2048), point counts proportional to the grid area,
    summing to about total_points."""
    areas = [grids[i % len(grids)][0] * grids[i % len(grids)][1] for i in range(n_scenes)]
    out = []
    for i in range(n_scenes):
        g = grids[i % len(grids)]
        target = int(total_points * areas[i] / sum(areas))
        out.append(floe_field(seed=seed + i, grid=g, h=h, target_points=target, dt=dt))
    return out
