"""Poisson-disk point clouds (the reference generator's sampling,
generator/2d/grainprocessor2d.cpp:40-56) against the properties that define
them: minimum distance, coverage, determinism, density and the argument checks.
Host code only (no GPU)."""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

import icempm_multi_amd as im
from icempm_multi_amd import capi, scenes


def test_minimum_distance_and_box():
    r = 0.1
    x, y = capi.poisson_disk(r, -1.0, 2.0, 4.0, 4.5, seed=11)
    assert len(x) > 700
    assert x.min() >= -1.0 and x.max() < 4.0 and y.min() >= 2.0 and y.max() < 4.5
    d, _ = cKDTree(np.stack([x, y], 1)).query(np.stack([x, y], 1), k=2)
    assert d[:, 1].min() >= r


def test_coverage_no_hole_of_radius_2r():
    # every active sample tried 30 candidates in its annulus [r, 2r) before
    # retiring, so almost every point of the box lies within 2r of a sample
    r = 0.05
    x, y = capi.poisson_disk(r, 0.0, 0.0, 2.0, 1.0, seed=5)
    g = np.stack(np.meshgrid(np.linspace(0, 2, 401), np.linspace(0, 1, 201)), -1).reshape(-1, 2)
    d, _ = cKDTree(np.stack([x, y], 1)).query(g)
    assert np.mean(d > 2 * r) < 1e-3
    assert d.max() < 3 * r


def test_deterministic_per_seed():
    a = capi.poisson_disk(0.07, 0, 0, 1.5, 1.0, seed=3)
    b = capi.poisson_disk(0.07, 0, 0, 1.5, 1.0, seed=3)
    c = capi.poisson_disk(0.07, 0, 0, 1.5, 1.0, seed=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert len(c[0]) != len(a[0]) or not np.array_equal(a[0], c[0])


def test_density_and_block_helper():
    r = 0.02
    x, _ = capi.poisson_disk(r, 0, 0, 4.0, 2.0, seed=9)
    fill = len(x) * r * r / 8.0
    assert 0.58 < fill < 0.66
    # poisson_block aims at 4 points per cell
    x, y = scenes.poisson_block(100, 40, 0.05, seed=2)
    assert abs(len(x) / (4 * 100 * 40) - 1) < 0.03


def test_argument_errors():
    with pytest.raises(im.IcempmError):
        capi.poisson_disk(0.0, 0, 0, 1, 1, seed=1)
    with pytest.raises(im.IcempmError):
        capi.poisson_disk(0.1, 0, 0, 0, 1, seed=1)
    x = np.empty(10)
    y = np.empty(10)
    n = C.c_longlong(0)
    rc = capi.lib().icempm_poisson_disk(0.05, 0.0, 0.0, 1.0, 1.0, 1, capi._dptr(x), capi._dptr(y), 10, C.byref(n))
    assert rc == -3   # ICEMPM_ERR_CAPACITY
