"""Edge cases of the point set through the C-ABI: no points, a single point,
a count that fills no wave or tile evenly, and a batch with an empty scene.
Stepping must run, advance the clock, and agree with the FP64 oracle."""
import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

FLAG_SETS = [0, im.STEP_GRAPH, im.STEP_UNFUSED, im.STEP_DIRECT_P2G]


def make(P, h, n, precision, cap=None):
    part = im.Partition(0, 0, precision)
    part.allocate(cap or h.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(h, 0, n)
    return part


def moving_block(seed):
    P, h, n = scenes.block_scene(nx=48, ny=12, grid=(96, 48), n_grains=40, seed=seed)
    h[6, :n] = 0.4
    h[7, :n] = -0.2 * np.sin(h[4, :n])
    return P, h, n


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED])
@pytest.mark.parametrize("flags", FLAG_SETS)
def test_no_points(precision, flags):
    P, h, n = moving_block(0x1CE0A01)
    part = make(P, h, 0, precision, cap=64)
    if not flags & im.STEP_DIRECT_P2G:
        part.sort_points()
    assert part.nPts_partition == 0
    part.step(12, flags)
    part.step(7, flags)
    assert part.error_code == 0
    t, step = part.clock
    assert step == 19 and t > 0
    out = np.zeros((16, 64))
    part.transfer_from_device(out, 0)
    assert np.all(out == 0)
    assert np.all(part.indenter_force() == 0)
    part.close()


@pytest.mark.parametrize("count", [1, 2, 63, 257, 1031])
@pytest.mark.parametrize("flags", FLAG_SETS)
def test_few_points_match_oracle(count, flags):
    """Counts below and between wave (64) and tile-iteration (256) multiples."""
    P, h, n = moving_block(0x1CE0A02)
    idx = np.random.default_rng(count).choice(n, size=count, replace=False)
    hs = np.zeros((16, count + 8))
    hs[:, :count] = h[:, idx]
    P.nPtsTotal = count
    ref = hs.copy()
    part = make(P, hs, count, im.PRECISION_FP64)
    if not flags & im.STEP_DIRECT_P2G:
        part.sort_points()
        O.sort(P, ref, count)
    part.step(9, flags)
    part.step(11, flags)
    t, step, _, err = O.substeps(P, ref, count, 20, 0.0, 0)
    assert err == 0 and part.error_code == 0
    out = np.zeros_like(hs)
    part.transfer_from_device(out, 0)
    for line in range(3, 16):
        sc = np.abs(ref[line, :count]).max() + 1e-30
        assert np.abs(out[line, :count] - ref[line, :count]).max() <= 1e-9 * sc + 1e-13, (count, flags, line)
    assert part.clock == (t, step)
    part.close()


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED])
def test_batch_with_an_empty_scene(precision):
    P1, h1, n1 = moving_block(0x1CE0A03)
    P2, h2, n2 = moving_block(0x1CE0A04)
    bt = im.Batch(0, precision)
    bt.add_scene(P1, h1, n1)
    bt.add_scene(P2, h2, 0)
    bt.add_scene(P2, h2, n2)
    bt.commit()
    bt.step(30, im.STEP_GRAPH | im.STEP_RESORT)
    bt.step(10, 0)
    assert bt.error_code == 0
    ref = h1.copy()
    O.sort(P1, ref, n1)
    O.substeps(P1, ref, n1, 40, 0.0, 0)
    out = np.zeros_like(h1)
    bt.transfer_from_device(0, out)
    tol = 1e-9 if precision == im.PRECISION_FP64 else 1e-4
    assert np.linalg.norm(out[4:6, :n1] - ref[4:6, :n1]) <= tol * np.linalg.norm(ref[4:6, :n1])
    empty = np.zeros_like(h2)
    bt.transfer_from_device(1, empty)
    assert np.all(empty == 0)
    bt.close()


def test_batch_commit_twice_is_refused():
    # a second commit must fail cleanly (the host copies of the scenes are
    # released by the first) and leave the committed batch usable
    P1, h1, n1 = moving_block(0x1CE0A05)
    bt = im.Batch(0, im.PRECISION_FP64)
    bt.add_scene(P1, h1, n1)
    bt.commit()
    with pytest.raises(im.IcempmError) as e:
        bt.commit()
    assert e.value.code == im.ERR_STATE and "already committed" in e.value.message
    with pytest.raises(im.IcempmError):
        bt.add_scene(P1, h1, n1)
    bt.step(10, 0)
    ref = h1.copy()
    O.sort(P1, ref, n1)
    O.substeps(P1, ref, n1, 10, 0.0, 0)
    out = np.zeros_like(h1)
    bt.transfer_from_device(0, out)
    assert np.linalg.norm(out[4:6, :n1] - ref[4:6, :n1]) <= 1e-9 * np.linalg.norm(ref[4:6, :n1])
    bt.close()


def test_negative_lengths_are_argument_errors():
    import ctypes as C
    P, h, n = moving_block(0x1CE0A06)
    part = make(P, h, n, im.PRECISION_FP64)
    part.sort_points()
    L = im.lib()
    buf = np.zeros(16)
    ibuf = np.zeros(16, dtype=np.int32)
    d = buf.ctypes.data_as(C.POINTER(C.c_double))
    i = ibuf.ctypes.data_as(C.POINTER(C.c_int32))
    assert L.icempm_partition_get_indenter_force(part._h, d, -1) == im.ERR_ARGUMENT
    assert L.icempm_partition_get_bins(part._h, i, None, -1, None, 0) == im.ERR_ARGUMENT
    assert L.icempm_partition_get_bins(part._h, None, None, 0, i, -5) == im.ERR_ARGUMENT
    assert L.icempm_partition_get_bins(part._h, i, i, 16, None, 0) == im.OK
    part.step(3, 0)
    assert part.error_code == 0
    part.close()


def test_sort_keys_of_points_beyond_the_grid():
    """The device key of a point beyond the grid is the key of the nearest
    cell the scene's cell table holds (columns -halo .. GridX+halo-1, rows
    0 .. GridY-1): the oracle key of the point projected there.  The
    reference's key (proxypoint2d.cpp:164-169) is not bounded and would
    index outside the table; such points also set the error byte when
    stepped.  Points inside the grid keep the oracle's key and order.  Only
    the sort runs here."""
    P, h, n = moving_block(0x1CE0A20)
    hh, gy, gx = P.GridHaloSize, P.GridY, P.GridXTotal
    cs = P.cellsize
    outside = [(-30 * cs, 5.3 * cs), ((gx + 30) * cs, 7.6 * cs), (20.2 * cs, -5 * cs), (31.7 * cs, (gy + 7) * cs),
               (-50 * cs, -50 * cs), ((gx + 50) * cs, (gy + 50) * cs), (-(hh + 3) * cs, 10 * cs)]
    m = n + len(outside)
    hs = np.zeros((16, m + 8))
    hs[:, :n] = h[:, :n]
    for k, (x, y) in enumerate(outside):
        hs[:, n + k] = h[:, k]
        hs[4, n + k], hs[5, n + k] = x, y
    P.nPtsTotal = m
    part = make(P, hs, m, im.PRECISION_FP64)
    part.sort_points()
    keys, perm, _ = part.get_bins(m, O.n_nodes(P))
    assert np.array_equal(np.sort(perm), np.arange(m)) and np.all(np.diff(keys) >= 0)
    # projection onto the table: column 0 is ix = -halo, reached by (int)(x/h + 0.5)
    # truncating toward zero at x = -(halo + 1) h; the last column at (GridX + halo - 1) h
    x_lo, x_hi = -(hh + 1) * cs, (gx + hh - 1) * cs
    y_lo, y_hi = 0.0, (gy - 1) * cs
    for k in range(m):
        x, y = hs[4, perm[k]], hs[5, perm[k]]
        if perm[k] < n:
            assert keys[k] == O.cell_key(P, x, y)
        else:
            xc = x if (x / cs + 0.5 >= -hh and x / cs + 0.5 < gx + hh) else min(max(x, x_lo), x_hi)
            yc = y if (y / cs + 0.5 >= 0 and y / cs + 0.5 < gy) else min(max(y, y_lo), y_hi)
            assert keys[k] == O.cell_key(P, xc, yc), (k, x, y, keys[k])
    # in-grid points: the oracle's stable order
    ref = hs.copy()
    _, rkeys, rperm = O.sort(P, ref, n)
    inside = perm[perm < n]
    assert np.array_equal(inside, rperm)
    part.close()
