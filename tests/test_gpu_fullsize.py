"""The 4.0M-point ice sheet (grid 4096 x 1024, h = 0.05 m, moving indenter) at
full size, through size-independent properties: the binning is a stable
counting sort (keys non-decreasing, permutation, ties in previous order,
counts = histogram) whose keys equal the oracle's on a seeded sample of
points; the grid mass equals N*m; the state stays finite with no error code.
The oracle runs only on the sample, so these finish in seconds."""
import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

SAMPLE = 20_000


@pytest.fixture(scope="module")
def sheet():
    return scenes.sheet_scene()


def check_bins(part, P, state, n, fixed=None):
    ncells = (P.GridXTotal + 2 * P.GridHaloSize) * P.GridY
    m = part.nPts_partition
    assert m == n
    keys, perm, counts = part.get_bins(m, ncells)
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(m))
    same = np.diff(keys) == 0
    assert np.all(np.diff(perm)[same] > 0), "ties must keep their previous order"
    assert np.array_equal(counts, np.bincount(keys + P.GridHaloSize * P.GridY, minlength=ncells))
    idx = np.random.default_rng(11).choice(m, size=min(SAMPLE, m), replace=False)
    for i in idx:
        j = perm[i]                       # previous index of sorted point i
        if fixed is None:
            k = O.cell_key(P, state[4, j], state[5, j])
        else:
            fb, ox, oy = fixed
            xq = int(np.rint((state[4, j] / P.cellsize + ox) * 2.0 ** fb))
            yq = int(np.rint((state[5, j] / P.cellsize + oy) * 2.0 ** fb))
            k = O.cell_key_fixed(xq, yq, fb, ox, oy, P.GridY)
        assert keys[i] == k, (i, j, keys[i], k)


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED])
def test_sheet_4m_binning_mass_and_state(sheet, precision):
    P, h, n = sheet
    assert abs(n - 4_000_000) < 40_000            # Poisson-disk cloud: 4 points per cell on average
    part = im.Partition(0, 0, precision)
    part.allocate(h.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(h, 0, n)
    part.sort_points()
    part.step(200, im.STEP_GRAPH)
    state = np.zeros_like(h)
    part.transfer_from_device(state, 0)        # positions before the re-sort
    assert part.error_code == 0
    assert np.all(np.isfinite(state[1:16, :n]))
    part.sort_points()
    fixed = part.position_map() if precision == im.PRECISION_MIXED else None
    check_bins(part, P, state, n, fixed)
    # mass: one P2G of the sorted state
    part.reset_grid()
    part.p2g()
    g = part.get_grid(0, O.n_nodes(P))
    # the mixed form carries the particle mass in FP32 (0.6125 kg is not exact
    # in binary); the FP64 grid sums then conserve that mass
    m = P.ParticleMass if precision == im.PRECISION_FP64 else float(np.float32(P.ParticleMass))
    M = n * m
    assert abs(g[0].sum() - M) <= 1e-9 * M
    # the indenter force accumulator stays finite
    f = part.indenter_force().reshape(-1, 2).sum(0)
    assert np.isfinite(f).all()
    part.close()
