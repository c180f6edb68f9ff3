"""Tile windows with a drift margin of 0 and of 1 cells (both window heights
are compiled in; the autotuner picks one): a block that translates by about
0.6 cells after the sort has many points outside their margin-0 windows, which
take the global path, and every configuration matches the FP64 oracle."""
import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

NSUB = 300


def translating_block():
    P, h, n = scenes.block_scene(nx=48, ny=12, grid=(96, 48), n_grains=40, seed=0x1CE0C01)
    h[6, :n] = 50.0 + 0.5 * np.sin(h[4, :n])     # m/s; 50 * 300 * 1e-5 = 0.15 m = 0.6 cells
    h[7, :n] = 20.0 + 0.3 * np.cos(h[5, :n])
    return P, h, n


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(8, 32), (4, 16), (8, 64)], ids=["8x32", "4x16", "8x64"])
@pytest.mark.parametrize("margin", [0, 1])
def test_window_margin_matches_oracle(precision, shape, margin):
    P, h, n = translating_block()
    ref0 = h.copy()
    O.sort(P, ref0, n)
    ref = ref0.copy()
    t, step, _f, err = O.substeps(P, ref, n, NSUB, 0.0, 0)
    assert err == 0
    assert np.abs(ref[4, :n] - ref0[4, :n]).mean() >= 0.5 * P.cellsize   # points left their sorted cells
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
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    if precision == im.PRECISION_FP64:
        for lines in ((4, 6), (6, 8), (8, 12), (12, 16)):
            sc = np.abs(ref[lines[0]:lines[1], :n]).max()
            assert np.abs(out[lines[0]:lines[1], :n] - ref[lines[0]:lines[1], :n]).max() <= 1e-9 * sc, lines
    else:
        assert rel(out[4:6, :n], ref[4:6, :n]) <= 1e-6
        assert rel(out[6:8, :n], ref[6:8, :n]) <= 1e-4


def test_batch_points_set_aside_match_oracle():
    """Three translating blocks of different speeds in one batch, margin-0
    windows: in every substep many points of every scene leave their window
    and go to the launch's list of points set aside, which its last workgroup
    handles (each entry carries its scene).  Every scene matches its own FP64
    oracle run."""
    bag = []
    for k, (vx, vy) in enumerate(((50.0, 20.0), (-40.0, 10.0), (30.0, -25.0))):
        P, h, n = scenes.block_scene(nx=40 + 8 * k, ny=10 + 2 * k, grid=(96, 48), n_grains=30, seed=0x1CE0C10 + k)
        h[6, :n] = vx + 0.5 * np.sin(h[4, :n])
        h[7, :n] = vy + 0.3 * np.cos(h[5, :n])
        bag.append((P, h, n))
    b = im.Batch(0, im.PRECISION_FP64)
    for P, h, n in bag:
        b.add_scene(P, h, n)
    b.commit()
    b.set_tile_shape(8, 32, 0)
    b.step(NSUB // 2, im.STEP_RESORT)
    b.step(NSUB - NSUB // 2, im.STEP_GRAPH)
    assert b.error_code == 0
    for i, (P, h, n) in enumerate(bag):
        ref = h.copy()
        O.sort(P, ref, n)
        _, _, _, err = O.substeps(P, ref, n, NSUB, 0.0, 0)
        assert err == 0
        out = np.zeros_like(h)
        b.transfer_from_device(i, out)
        for lines in ((4, 6), (6, 8), (8, 12), (12, 16)):
            sc = np.abs(ref[lines[0]:lines[1], :n]).max()
            d = np.abs(out[lines[0]:lines[1], :n] - ref[lines[0]:lines[1], :n]).max()
            assert d <= 1e-9 * sc, (i, lines, d, sc)
    b.close()
