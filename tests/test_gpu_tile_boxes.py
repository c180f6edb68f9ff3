"""Tile windows over the box of each tile's occupied cells: a sparse scene of
scattered point clusters (one to three cells each, about 12% of the cells)
gives tiles whose boxes are single cells, strips along every tile edge and
whole tiles.  Stepped with two tile shapes at both drift margins (points that
leave their box's window take the global path), every configuration matches
the FP64 oracle; the mixed form matches it to its position bar."""
import math

import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

NSUB = 200
H = 0.25


def sparse_scene(seed=0x1CE0D01):
    rng = np.random.default_rng(seed)
    grid = (96, 48)
    tan_phi = math.tan(30.0 * math.pi / 180)
    P = scenes.make_params(GridX=grid[0], GridY=grid[1], GridXDimension=grid[0] * H, InitialTimeStep=1e-5,
                           YoungsModulus=9e9, PoissonsRatio=0.3, Density=980.0, DP_phi=30.0,
                           DP_threshold_p=1e5 / tan_phi, SetupType=2, IndVelocity=0.0, GridHaloSize=8,
                           nPartitions=1)
    xs, ys = [], []
    cells = set()
    while len(cells) < 0.12 * (grid[0] - 24) * (grid[1] - 12):
        cx, cy = int(rng.integers(4, grid[0] - 20)), int(rng.integers(4, grid[1] - 8))
        w, hgt = int(rng.integers(1, 4)), int(rng.integers(1, 3))
        for i in range(w):
            for j in range(hgt):
                cells.add((cx + i, cy + j))
    for k, (cx, cy) in enumerate(sorted(cells)):
        x, y = scenes.jittered_block(1, 1, H, seed + k, x0=cx * H, y0=cy * H)
        xs.append(x)
        ys.append(y)
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    g = np.arange(len(x), dtype=np.uint64) % 97
    h, n = scenes.load_raw_points(P, x, y, g, len(cells) * H * H)
    u = rng.uniform(-1.0, 1.0, (2, n))
    h[6, :n] = 40.0 * u[0]          # up to 40 m/s: 0.08 m = 0.32 cells over the run
    h[7, :n] = 40.0 * u[1]
    return P, h, n


def boxes(P, h, n, tx, ty):
    """(box columns, box rows, offsets in the tile) of every non-empty tile."""
    col = (h[4, :n] / P.cellsize + 0.5).astype(np.int64)
    row = (h[5, :n] / P.cellsize + 0.5).astype(np.int64)
    t = (col // tx) * 10_000 + row // ty
    out = []
    for k in np.unique(t):
        m = t == k
        c, r = col[m] % tx, row[m] % ty
        out.append((c.max() - c.min() + 1, r.max() - r.min() + 1, c.min(), c.max(), r.min(), r.max()))
    return out


def test_sparse_scene_has_every_kind_of_box():
    P, h, n = sparse_scene()
    for tx, ty in ((8, 32), (4, 16)):
        b = boxes(P, h, n, tx, ty)
        assert any(bw == 1 and bh == 1 for bw, bh, *_ in b)                 # single cells
        assert any(c0 == 0 for _, _, c0, *_ in b) and any(c1 == tx - 1 for _, _, _, c1, _, _ in b)
        assert any(r0 == 0 for *_, r0, _ in b) and any(r1 == ty - 1 for *_, r1 in b)
        assert any(bw < tx and bh < ty for bw, bh, *_ in b)                 # boxes smaller than the tile


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(8, 32), (4, 16)], ids=["8x32", "4x16"])
@pytest.mark.parametrize("margin", [0, 1])
def test_sparse_tiles_match_oracle(precision, shape, margin):
    P, h, n = sparse_scene()
    ref = h.copy()
    O.sort(P, ref, n)
    t, step, _f, err = O.substeps(P, ref, n, NSUB, 0.0, 0)
    assert err == 0
    part = im.Partition(0, 0, precision)
    part.allocate(h.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.set_tile_shape(shape[0], shape[1], margin)
    part.transfer_points_from_soa_to_device(h, 0, n)
    part.sort_points()
    part.step(NSUB // 2, 0)
    part.step(NSUB - NSUB // 2, im.STEP_GRAPH)
    out = np.zeros_like(h)
    part.transfer_from_device(out, 0)
    assert part.error_code == 0
    assert part.clock[1] == step
    part.close()
    if precision == im.PRECISION_FP64:
        for lines in ((4, 6), (6, 8), (8, 12), (12, 16)):
            sc = np.abs(ref[lines[0]:lines[1], :n]).max()
            assert np.abs(out[lines[0]:lines[1], :n] - ref[lines[0]:lines[1], :n]).max() <= 1e-9 * sc, lines
    else:
        rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel(out[4:6, :n], ref[4:6, :n]) <= 1e-6
        assert rel(out[6:8, :n], ref[6:8, :n]) <= 1e-4
