"""Elastic P and S fronts in an ice block stepped on the GPU through the C-ABI,
in both precisions: their speeds are the ones the linearised Kirchhoff stress
fixes, to the bar the oracle is pinned to (tests/_waves.py)."""
import numpy as np
import pytest

import _oracle as O
import _waves as W
import icempm_multi_amd as im

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("comp", [0, 1], ids=["P", "S"])
def test_elastic_wave_speeds(precision, comp):
    P, hs, n = W.block()
    O.sort(P, hs, n)          # the device sort then keeps this order
    probes, nsub = W.start(P, hs, n, comp)
    part = im.Partition(0, 0, precision)
    part.allocate(hs.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(hs, 0, n)
    part.sort_points()
    out = np.zeros_like(hs)
    times, series = [], [[] for _ in probes]
    for k in range(nsub // W.CHUNK):
        part.step(W.CHUNK, im.STEP_GRAPH)
        part.transfer_from_device(out, 0)
        if k == 0:   # same points in the same order (they have moved < 1 um)
            assert np.abs(out[4:6, :n] - hs[4:6, :n]).max() < 1e-5
        times.append(part.clock[0])
        for s, m in zip(series, probes):
            s.append(out[6 + comp, :n][m].mean())
    assert part.error_code == 0
    part.close()
    ratio = W.measured_speed(times, series) / W.expected_speed(P, comp)
    assert 0.97 <= ratio <= 1.01, ratio
    for s in series:
        assert abs(s[-1] - W.V / 2) <= 0.1 * W.V / 2, s[-1]
