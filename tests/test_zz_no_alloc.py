"""icempm_substeps_noalloc, the entry point that allocates no device memory
(no counterpart in the reference), against icempm_partition_step with
ICEMPM_STEP_DIRECT_P2G, whose kernels it runs.

Two calls in a row on one GPU, the second with fewer points and a smaller
grid.  Each must match the partition's result on the same input: every field
within 1e-10 of its scale (the order of the atomic sums differs), identical
utility words, the same error byte, and the same indenter force bins to 1e-8.
The refusals (grid beyond the static array, bad sizes) happen on the host
before the GPU is touched."""
import ctypes as C

import numpy as np
import pytest

import icempm_multi_amd as im
from icempm_multi_amd import capi, scenes

FIELDS = {"x": (4, 6), "v": (6, 8), "F": (8, 12), "C": (12, 16), "Jp": (3, 4), "PQ": (1, 3)}


def with_motion(P, h, n, seed):
    r = np.random.default_rng(seed)
    h[6, :n] = 0.4 + 0.2 * np.sin(3 * h[4, :n])
    h[7, :n] = -0.3 * np.cos(2 * h[5, :n])
    h[8:12, :n] += 2e-4 * r.normal(size=(4, n))
    h[12:16, :n] = 0.5 * r.normal(size=(4, n))
    return P, h, n


def hip():
    """The HIP runtime libicempm.so itself links (device memory for the
    caller's buffers; a second runtime in the process, such as torch's,
    would not see the GPU)."""
    L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipFree.argtypes = [C.c_void_p]
    return L


class DeviceArray:
    """A device copy of a numpy array, copied back by get()."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.a.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0      # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0          # device to host
        return out

    def free(self):
        hip().hipFree(self.ptr)


def noalloc(P, h, n, nsub):
    pitch = h.shape[1]
    d = DeviceArray(h)
    err = DeviceArray(np.full(1, 255, dtype=np.uint32))
    force = DeviceArray(np.full(2 * P.n_indenter_subdivisions, -1.0))
    print(f"noalloc: {n} points, grid {P.GridXTotal}x{P.GridY}, {nsub} substeps", flush=True)
    rc = capi.lib().icempm_substeps_noalloc(0, C.byref(P), C.sizeof(im.SimParams), d.ptr, pitch, n, nsub, None,
                                            err.ptr, force.ptr)
    assert rc == 0, capi.lib().icempm_last_error()
    assert hip().hipDeviceSynchronize() == 0
    out = d.get(), int(err.get()[0]), force.get()
    for x in (d, err, force):
        x.free()
    return out


def partition(P, h, n, nsub):
    part = im.Partition(0, 0, im.PRECISION_FP64)
    part.allocate(h.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(h, 0, n)
    part.reset_indenter_force_accumulator()
    part.step(nsub, im.STEP_DIRECT_P2G)
    out = np.zeros_like(h)
    part.transfer_from_device(out, 0)
    res = out, part.error_code, part.indenter_force(P.n_indenter_subdivisions)
    part.close()
    return res


def assert_same(a, b, n, what):
    for name, (lo, hi) in FIELDS.items():
        s = max(np.abs(b[lo:hi, :n]).max(), 1e-300)
        d = np.abs(a[lo:hi, :n] - b[lo:hi, :n]).max()
        assert d <= 1e-10 * s, f"{what} {name}: max diff {d:.3e} scale {s:.3e}"
    assert np.array_equal(a[0, :n].view(np.int64), b[0, :n].view(np.int64)), f"{what}: utility words"


@pytest.mark.gpu
def test_noalloc_matches_partition_step_twice():
    # first call: a sheet with the moving indenter (force bins filled);
    # second: a smaller block on a smaller grid
    first = with_motion(*scenes.sheet_scene(seed=0x1CE0D01, nx=200, ny=20, grid=(256, 64), h=0.05,
                                            ind_diameter=0.5, ind_depth=0.15, n_grains=40, sampler="jittered"), 5)
    second = with_motion(*scenes.block_scene(seed=0x1CE0D02, nx=30, ny=8, grid=(64, 32), n_grains=20), 6)
    assert second[2] < first[2]
    for (P, h, n), nsub in ((first, 12), (second, 9)):
        got, err, force = noalloc(P, h, n, nsub)
        ref, err_ref, force_ref = partition(P, h, n, nsub)
        what = f"{n} points, grid {P.GridXTotal}x{P.GridY}"
        assert err == err_ref == 0, (what, err, err_ref)
        assert_same(got, ref, n, what)
        # the points beyond n are left alone
        assert np.array_equal(got[:, n:], h[:, n:])
        assert np.allclose(force, force_ref, rtol=1e-8, atol=1e-8 * max(np.abs(force_ref).max(), 1e-300)), what
    assert np.abs(force).max() == 0 and np.abs(noalloc(*first, 3)[2]).max() > 0   # block: no indenter; sheet: contact


def test_noalloc_refusals_on_the_host():
    """Refused before any GPU call, so this runs without a GPU too."""
    L = capi.lib()
    P, h, n = scenes.block_scene(seed=0x1CE0D03, nx=8, ny=4, grid=(64, 32), n_grains=4)
    sz = C.sizeof(im.SimParams)
    big = im.params_derive(im.params_reset())
    big.GridXTotal, big.GridY, big.GridHaloSize = 2048, 1024, 8
    big.n_indenter_subdivisions = P.n_indenter_subdivisions
    assert (2048 + 16) * 1024 > (1 << 20)          # ICEMPM_NOALLOC_MAX_NODES
    assert L.icempm_substeps_noalloc(0, C.byref(big), sz, None, 64, 0, 1, None, None, None) == im.ERR_CAPACITY
    assert "at most" in L.icempm_last_error().decode()
    assert L.icempm_substeps_noalloc(0, C.byref(P), sz - 8, None, 64, 0, 1, None, None, None) == im.ERR_ARGUMENT
    assert L.icempm_substeps_noalloc(0, C.byref(P), sz, None, 64, 65, 1, None, None, None) == im.ERR_ARGUMENT
    assert L.icempm_substeps_noalloc(0, C.byref(P), sz, None, 64, 10, 1, None, None, None) == im.ERR_ARGUMENT
    assert L.icempm_substeps_noalloc(0, None, sz, None, 64, 0, 1, None, None, None) == im.ERR_ARGUMENT
