"""Misuse of the C-ABI is refused with the ICEMPM_ERR_* code for it and a
message, as the reference throws std::runtime_error (gpu_partition.cu,
gpu_implementation5.cpp), and the refused call leaves the object usable."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu


def small():
    return scenes.block_scene(nx=48, ny=12, grid=(96, 48), n_grains=40, seed=0x1CE0B01)


def refused(code, fn, *args):
    with pytest.raises(im.IcempmError) as e:
        fn(*args)
    assert e.value.code == code, (e.value.code, e.value.message)
    assert e.value.message


def steps_like_oracle(part, P, h, n, tol, nsub=10):
    ref = h.copy()
    O.sort(P, ref, n)
    O.substeps(P, ref, n, nsub, 0.0, 0)
    part.step(nsub, 0)
    out = np.zeros_like(h)
    part.transfer_from_device(out, 0)
    assert part.error_code == 0
    assert np.linalg.norm(out[4:6, :n] - ref[4:6, :n]) <= tol * np.linalg.norm(ref[4:6, :n])


def test_partition_create_arguments():
    refused(im.ERR_ARGUMENT, im.Partition, 0, 0, 7)           # unknown precision
    refused(im.ERR_ARGUMENT, im.Partition, 1 << 20, 0, im.PRECISION_FP64)
    refused(im.ERR_ARGUMENT, im.Batch, -1, im.PRECISION_MIXED)


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED])
def test_partition_call_order_and_sizes(precision):
    P, h, n = small()
    part = im.Partition(0, 0, precision)
    refused(im.ERR_STATE, part.update_constants, P)             # before allocate
    refused(im.ERR_ARGUMENT, part.allocate, 0, P.GridXTotal)
    part.allocate(h.shape[1], P.GridXTotal)
    refused(im.ERR_ARGUMENT, part.allocate, h.shape[1], P.GridXTotal)   # twice
    bad = P.copy()
    bad.cellsize = 0.0                                          # not derived / out of range
    refused(im.ERR_ARGUMENT, part.update_constants, bad)
    part.update_constants(P)
    refused(im.ERR_ARGUMENT, part.set_grid_span, 8, P.GridXTotal - 8)    # one partition spans the grid
    part.set_grid_span(0, P.GridXTotal)
    refused(im.ERR_CAPACITY, part.transfer_points_from_soa_to_device, h, 0, h.shape[1] + 1)
    refused(im.ERR_CAPACITY, part.transfer_points_from_soa_to_device, h, h.shape[1] - 10, 20)
    part.transfer_points_from_soa_to_device(h, 0, n)
    refused(im.ERR_STATE, part.get_bins, n, 0)                  # no sort since the upload
    small_soa = np.zeros((16, n - 1))
    refused(im.ERR_CAPACITY, part.transfer_from_device, small_soa, 0)
    refused(im.ERR_ARGUMENT, part.set_tile_shape, 0, 32)
    refused(im.ERR_ARGUMENT, part.set_tile_shape, 33, 32)
    refused(im.ERR_ARGUMENT, part.set_tile_shape, 32, 1024)    # window over the LDS budget
    refused(im.ERR_ARGUMENT, part.set_lane_map, 3)
    refused(im.ERR_ARGUMENT, part.set_tile_schedule, 2)
    nodes = (P.GridXTotal + 2 * P.GridHaloSize) * P.GridY
    refused(im.ERR_ARGUMENT, part.get_grid, 3, nodes)      # buffers 0, 1, 2
    refused(im.ERR_ARGUMENT, part.get_grid, 0, nodes + 1)
    part.sort_points()
    steps_like_oracle(part, P, h, n, 1e-9 if precision == im.PRECISION_FP64 else 1e-6)
    part.close()


def test_params_size_mismatch_is_refused():
    P, h, n = small()
    part = im.Partition(0, 0, im.PRECISION_FP64)
    part.allocate(h.shape[1], P.GridXTotal)
    L = im.lib()
    rc = L.icempm_partition_update_constants(part._h, C.byref(P), C.sizeof(im.SimParams) - 8)
    assert rc == im.ERR_ARGUMENT and b"size" in L.icempm_last_error()
    part.update_constants(P)
    part.close()


def test_batch_call_order_and_ids():
    P1, h1, n1 = small()
    bt = im.Batch(0, im.PRECISION_FP64)
    refused(im.ERR_STATE, bt.commit)                            # no scene
    refused(im.ERR_ARGUMENT, bt.add_scene, P1, h1, h1.shape[1] + 1)
    bt.add_scene(P1, h1, n1)
    refused(im.ERR_STATE, bt.step, 5)                           # not committed
    refused(im.ERR_STATE, bt.autotune)
    bt.commit()
    out = np.zeros_like(h1)
    refused(im.ERR_ARGUMENT, bt.transfer_from_device, 1, out)
    refused(im.ERR_ARGUMENT, bt.transfer_from_device, -1, out)
    refused(im.ERR_ARGUMENT, bt.indenter_force, 3)
    bt.step(10, 0)
    ref = h1.copy()
    O.sort(P1, ref, n1)
    O.substeps(P1, ref, n1, 10, 0.0, 0)
    bt.transfer_from_device(0, out)
    assert bt.error_code == 0
    assert np.linalg.norm(out[4:6, :n1] - ref[4:6, :n1]) <= 1e-9 * np.linalg.norm(ref[4:6, :n1])
    bt.close()
