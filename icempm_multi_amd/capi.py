"""ctypes binding of libicempm.so (include/icempm.h).

``Partition`` mirrors the reference's ``GPU_Partition`` method names
(simulation/gpu_partition.h:79-165); ``Batch`` is the many-scene form.  Host
particle data uses the reference's HostSideSOA layout: a float64 array of shape
(16, capacity), line ``i`` holding field ``i`` of every point
(simulation/parameters_sim.h:27-35); line 0 holds int64 bit patterns.

Errors raise ``IcempmError``, as the reference throws ``std::runtime_error``.
There is no fallback: without the built library every call fails loudly.
"""
import ctypes as C
import os

import numpy as np

from .params import SimParams, N_PTS_ARRAYS, PRECISION_FP64

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libicempm.so")

_lib = None


class IcempmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"icempm error {code}: {message}")
        self.code = code
        self.message = message


_P = C.POINTER(SimParams)
_D = C.POINTER(C.c_double)
_VP = C.c_void_p

_SIGS = {
    "icempm_version": (C.c_char_p, []),
    "icempm_last_error": (C.c_char_p, []),
    "icempm_params_reset": (C.c_int, [_P]),
    "icempm_params_parse_json": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_size_t]),
    "icempm_params_derive": (C.c_int, [_P]),
    "icempm_prepare_raw_points": (C.c_int, [_P, _D, C.c_uint, C.c_int, C.c_float, C.POINTER(C.c_int)]),
    "icempm_partition_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_VP)]),
    "icempm_partition_destroy": (C.c_int, [_VP]),
    "icempm_partition_allocate": (C.c_int, [_VP, C.c_int, C.c_int]),
    "icempm_partition_update_constants": (C.c_int, [_VP, _P, C.c_size_t]),
    "icempm_partition_set_grid_span": (C.c_int, [_VP, C.c_int, C.c_int]),
    "icempm_partition_transfer_points_from_soa_to_device": (C.c_int, [_VP, _D, C.c_uint, C.c_int, C.c_int]),
    "icempm_partition_transfer_from_device": (C.c_int, [_VP, _D, C.c_uint, C.c_int]),
    "icempm_partition_reset_grid": (C.c_int, [_VP]),
    "icempm_partition_reset_indenter_force_accumulator": (C.c_int, [_VP]),
    "icempm_partition_p2g": (C.c_int, [_VP]),
    "icempm_partition_update_nodes": (C.c_int, [_VP, C.c_double, C.c_double, C.c_double]),
    "icempm_partition_g2p": (C.c_int, [_VP, C.c_int]),
    "icempm_partition_synchronize": (C.c_int, [_VP]),
    "icempm_partition_get_indenter_force": (C.c_int, [_VP, _D, C.c_int]),
    "icempm_partition_get_error_code": (C.c_int, [_VP, C.POINTER(C.c_uint8)]),
    "icempm_partition_get_point_count": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "icempm_partition_get_timings": (C.c_int, [_VP, C.POINTER(C.c_float), C.c_int]),
    "icempm_partition_get_kernel_stats": (C.c_int, [_VP, _D, C.POINTER(C.c_longlong), C.c_int]),
    "icempm_partition_step": (C.c_int, [_VP, C.c_int, C.c_int]),
    "icempm_partition_get_clock": (C.c_int, [_VP, _D, C.POINTER(C.c_int)]),
    "icempm_partition_sort_points": (C.c_int, [_VP]),
    "icempm_partition_get_bins": (C.c_int, [_VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int,
                                            C.POINTER(C.c_int32), C.c_int]),
    "icempm_partition_get_grid": (C.c_int, [_VP, C.c_int, _D, C.c_longlong]),
    "icempm_partition_get_position_map": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                    C.POINTER(C.c_int)]),
    "icempm_partition_set_tile_shape": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    "icempm_partition_set_lane_map": (C.c_int, [_VP, C.c_int]),
    "icempm_partition_set_tile_schedule": (C.c_int, [_VP, C.c_int]),
    "icempm_partition_autotune": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    "icempm_partition_get_tile_config": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                   C.POINTER(C.c_int)]),
    "icempm_partition_get_graph_stats": (C.c_int, [_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "icempm_batch_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_VP)]),
    "icempm_batch_destroy": (C.c_int, [_VP]),
    "icempm_batch_add_scene": (C.c_int, [_VP, _P, C.c_size_t, _D, C.c_uint, C.c_int, C.POINTER(C.c_int)]),
    "icempm_batch_commit": (C.c_int, [_VP]),
    "icempm_batch_step": (C.c_int, [_VP, C.c_int, C.c_int]),
    "icempm_batch_transfer_from_device": (C.c_int, [_VP, C.c_int, _D, C.c_uint]),
    "icempm_batch_get_indenter_force": (C.c_int, [_VP, C.c_int, _D, C.c_int]),
    "icempm_batch_get_error_code": (C.c_int, [_VP, C.POINTER(C.c_uint8)]),
    "icempm_batch_get_scene_point_count": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int)]),
    "icempm_batch_get_bins": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int,
                                        C.POINTER(C.c_int32), C.c_int]),
    "icempm_batch_get_grid": (C.c_int, [_VP, C.c_int, C.c_int, _D, C.c_longlong]),
    "icempm_batch_get_position_map": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_int)]),
    "icempm_batch_get_scene_count": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "icempm_batch_last_step_ms": (C.c_int, [_VP, C.POINTER(C.c_float)]),
    "icempm_batch_get_kernel_stats": (C.c_int, [_VP, _D, C.POINTER(C.c_longlong), C.c_int]),
    "icempm_batch_reset_kernel_stats": (C.c_int, [_VP]),
    "icempm_poisson_disk": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, _D, _D,
                                      C.c_longlong, C.POINTER(C.c_longlong)]),
    "icempm_batch_autotune": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    "icempm_tile_candidates": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    "icempm_substeps_noalloc": (C.c_int, [C.c_int, _P, C.c_size_t, C.c_void_p, C.c_uint, C.c_int, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
    "icempm_batch_get_tile_config": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int)]),
    "icempm_batch_set_tile_shape": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    "icempm_batch_get_graph_stats": (C.c_int, [_VP, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)

KERNEL_CLASSES = ("clock", "zero", "p2g", "fused_g2p_p2g", "g2p", "update", "sort")


def lib():
    """Loads libicempm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise IcempmError(rc, lib().icempm_last_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(_D)


def _soa(hssoa):
    if not (isinstance(hssoa, np.ndarray) and hssoa.dtype == np.float64 and hssoa.ndim == 2
            and hssoa.shape[0] == N_PTS_ARRAYS and hssoa.flags.c_contiguous):
        raise ValueError("HostSideSOA buffer must be a C-contiguous float64 array of shape (16, capacity)")
    return hssoa


# ---------------------------------------------------------------- parameters

def params_reset():
    """icy::SimParams::Reset (parameters_sim.cpp:5-60)."""
    P = SimParams()
    _check(lib().icempm_params_reset(C.byref(P)))
    return P


def params_parse_json(text, P=None):
    """icy::SimParams::ParseFile on a JSON text; returns (params, InputRawPoints)."""
    if P is None:
        P = params_reset()
    buf = C.create_string_buffer(4096)
    _check(lib().icempm_params_parse_json(C.byref(P), text.encode(), buf, 4096))
    return P, buf.value.decode()


def params_derive(P):
    """ComputeCamClayParams2 + ComputeHelperVariables (parameters_sim.cpp:128-165)."""
    _check(lib().icempm_params_derive(C.byref(P)))
    return P


def prepare_raw_points(P, hssoa, n_points, volume):
    """SnapshotManager::LoadRawPoints after the file read (snapshotmanager.cpp:57-110)."""
    _soa(hssoa)
    kept = C.c_int(0)
    _check(lib().icempm_prepare_raw_points(C.byref(P), _dptr(hssoa), hssoa.shape[1], n_points,
                                           C.c_float(volume), C.byref(kept)))
    return kept.value


def poisson_disk(r, x0, y0, x1, y1, seed):
    """Poisson-disk sample of [x0, x1) x [y0, y1) with minimum distance r
    (generator/2d/grainprocessor2d.cpp:40-56, Bridson's algorithm); returns x, y."""
    # at most one point per disk of radius r/2: hexagonal packing bounds the count
    # (the library rejects r <= 0 and empty boxes)
    ok = r > 0 and x1 > x0 and y1 > y0
    cap = int(1.2 * (x1 - x0 + r) * (y1 - y0 + r) / (r * r)) + 16 if ok else 16
    x = np.empty(cap)
    y = np.empty(cap)
    n = C.c_longlong(0)
    _check(lib().icempm_poisson_disk(r, x0, y0, x1, y1, seed & 0xFFFFFFFFFFFFFFFF, _dptr(x), _dptr(y), cap,
                                     C.byref(n)))
    return x[:n.value].copy(), y[:n.value].copy()


def tile_candidates():
    """The autotuner's candidate shapes [(tx, ty, lane_spread), ...]; autotune()
    reports them at margin 0, then at margin 1, in this order."""
    k = lib().icempm_tile_candidates(None, None, None, 0)
    tx, ty, ls = (C.c_int * k)(), (C.c_int * k)(), (C.c_int * k)()
    lib().icempm_tile_candidates(tx, ty, ls, k)
    return list(zip(tx, ty, ls))


# ---------------------------------------------------------------- partition

class Partition:
    """One device partition (GPU_Partition): points and grid on one GPU."""

    def __init__(self, device=0, partition_id=0, precision=PRECISION_FP64):
        h = _VP()
        _check(lib().icempm_partition_create(device, partition_id, precision, C.byref(h)))
        self._h = h
        self.precision = precision
        self.PartitionID = partition_id
        self.Device = device

    def close(self):
        if self._h:
            lib().icempm_partition_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # GPU_Partition methods
    def allocate(self, n_points_capacity, grid_x_capacity):
        _check(lib().icempm_partition_allocate(self._h, n_points_capacity, grid_x_capacity))

    def update_constants(self, prms):
        _check(lib().icempm_partition_update_constants(self._h, C.byref(prms), C.sizeof(SimParams)))

    def set_grid_span(self, gridx_offset, gridx_partition):
        _check(lib().icempm_partition_set_grid_span(self._h, gridx_offset, gridx_partition))

    def transfer_points_from_soa_to_device(self, hssoa, point_idx_offset, n_points):
        _soa(hssoa)
        _check(lib().icempm_partition_transfer_points_from_soa_to_device(
            self._h, _dptr(hssoa), hssoa.shape[1], point_idx_offset, n_points))

    def transfer_from_device(self, hssoa, point_idx_offset=0):
        _soa(hssoa)
        _check(lib().icempm_partition_transfer_from_device(self._h, _dptr(hssoa), hssoa.shape[1], point_idx_offset))

    def reset_grid(self):
        _check(lib().icempm_partition_reset_grid(self._h))

    def reset_indenter_force_accumulator(self):
        _check(lib().icempm_partition_reset_indenter_force_accumulator(self._h))

    def p2g(self):
        _check(lib().icempm_partition_p2g(self._h))

    def update_nodes(self, simulation_time, indenter_x, indenter_y):
        _check(lib().icempm_partition_update_nodes(self._h, simulation_time, indenter_x, indenter_y))

    def g2p(self, recordPQ):
        _check(lib().icempm_partition_g2p(self._h, int(bool(recordPQ))))

    def synchronize(self):
        _check(lib().icempm_partition_synchronize(self._h))

    # beyond the reference
    def step(self, n_substeps, flags=0):
        _check(lib().icempm_partition_step(self._h, n_substeps, flags))

    def sort_points(self):
        _check(lib().icempm_partition_sort_points(self._h))

    def set_tile_shape(self, tx, ty, margin=1):
        _check(lib().icempm_partition_set_tile_shape(self._h, tx, ty, margin))

    def set_lane_map(self, lane_spread):
        _check(lib().icempm_partition_set_lane_map(self._h, lane_spread))

    def set_tile_schedule(self, dynamic):
        _check(lib().icempm_partition_set_tile_schedule(self._h, dynamic))

    def autotune(self, n_trials=5):
        """Returns (chosen index, ms per substep launch of the candidates of
        tile_candidates() at margin 0 then margin 1); tile_config() gives the
        shape kept."""
        k = 2 * len(tile_candidates())
        out = (C.c_float * k)()
        chosen = C.c_int(-1)
        _check(lib().icempm_partition_autotune(self._h, n_trials, out, k, C.byref(chosen)))
        return chosen.value, list(out)

    def tile_config(self):
        """{tx, ty, margin, lane_spread} of the binned kernels."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().icempm_partition_get_tile_config(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("tx", "ty", "margin", "lane_spread"), (x.value for x in v)))

    def graph_stats(self):
        """(graphs captured since creation, graphs cached now)."""
        n, c = C.c_longlong(), C.c_int()
        _check(lib().icempm_partition_get_graph_stats(self._h, C.byref(n), C.byref(c)))
        return n.value, c.value

    def indenter_force(self, n_subdivisions=100):
        out = np.zeros(2 * n_subdivisions)
        _check(lib().icempm_partition_get_indenter_force(self._h, _dptr(out), out.size))
        return out

    @property
    def error_code(self):
        c = C.c_uint8(0)
        _check(lib().icempm_partition_get_error_code(self._h, C.byref(c)))
        return c.value

    @property
    def nPts_partition(self):
        n, d = C.c_int(), C.c_int()
        _check(lib().icempm_partition_get_point_count(self._h, C.byref(n), C.byref(d)))
        return n.value

    @property
    def clock(self):
        t, s = C.c_double(), C.c_int()
        _check(lib().icempm_partition_get_clock(self._h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def timings(self):
        out = (C.c_float * 8)()
        _check(lib().icempm_partition_get_timings(self._h, out, 8))
        return list(out)

    def kernel_stats(self):
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = (C.c_longlong * len(KERNEL_CLASSES))()
        _check(lib().icempm_partition_get_kernel_stats(self._h, _dptr(ms), cnt, len(KERNEL_CLASSES)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def get_grid(self, buf, n_nodes):
        out = np.zeros((3, n_nodes))
        _check(lib().icempm_partition_get_grid(self._h, buf, _dptr(out), n_nodes))
        return out

    def position_map(self):
        fb, ox, oy = C.c_int(), C.c_int(), C.c_int()
        _check(lib().icempm_partition_get_position_map(self._h, C.byref(fb), C.byref(ox), C.byref(oy)))
        return fb.value, ox.value, oy.value

    def get_bins(self, n, n_cells):
        keys = np.zeros(max(n, 1), dtype=np.int32)
        perm = np.zeros(max(n, 1), dtype=np.int32)
        counts = np.zeros(max(n_cells, 1), dtype=np.int32)
        i32 = C.POINTER(C.c_int32)
        _check(lib().icempm_partition_get_bins(self._h, keys.ctypes.data_as(i32), perm.ctypes.data_as(i32), n,
                                               counts.ctypes.data_as(i32), n_cells))
        return keys[:n], perm[:n], counts[:n_cells]


# ---------------------------------------------------------------- batch

class Batch:
    """Many independent scenes stepped by one persistent launch per phase."""

    def __init__(self, device=0, precision=PRECISION_FP64):
        h = _VP()
        _check(lib().icempm_batch_create(device, precision, C.byref(h)))
        self._h = h
        self.precision = precision
        self.scenes = []

    def close(self):
        if self._h:
            lib().icempm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, prms, hssoa, n_points):
        _soa(hssoa)
        sid = C.c_int(-1)
        _check(lib().icempm_batch_add_scene(self._h, C.byref(prms), C.sizeof(SimParams), _dptr(hssoa),
                                            hssoa.shape[1], n_points, C.byref(sid)))
        self.scenes.append((prms.copy(), n_points))
        return sid.value

    def commit(self):
        _check(lib().icempm_batch_commit(self._h))

    def step(self, n_substeps, flags=0):
        _check(lib().icempm_batch_step(self._h, n_substeps, flags))

    def transfer_from_device(self, scene_id, hssoa):
        _soa(hssoa)
        _check(lib().icempm_batch_transfer_from_device(self._h, scene_id, _dptr(hssoa), hssoa.shape[1]))

    def indenter_force(self, scene_id, n_subdivisions=100):
        out = np.zeros(2 * n_subdivisions)
        _check(lib().icempm_batch_get_indenter_force(self._h, scene_id, _dptr(out), out.size))
        return out

    @property
    def error_code(self):
        c = C.c_uint8(0)
        _check(lib().icempm_batch_get_error_code(self._h, C.byref(c)))
        return c.value

    def counts(self):
        n, t = C.c_int(), C.c_longlong()
        _check(lib().icempm_batch_get_scene_count(self._h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def scene_points(self, scene_id):
        """Points scene_id holds (after the last compaction)."""
        n = C.c_int()
        _check(lib().icempm_batch_get_scene_point_count(self._h, scene_id, C.byref(n)))
        return n.value

    def get_bins(self, scene_id, n, n_cells):
        """(sorted keys, previous index of each sorted point, cell counts) of scene_id."""
        keys = np.zeros(max(n, 1), dtype=np.int32)
        perm = np.zeros(max(n, 1), dtype=np.int32)
        counts = np.zeros(max(n_cells, 1), dtype=np.int32)
        i32 = C.POINTER(C.c_int32)
        _check(lib().icempm_batch_get_bins(self._h, scene_id, keys.ctypes.data_as(i32), perm.ctypes.data_as(i32), n,
                                           counts.ctypes.data_as(i32), n_cells))
        return keys[:n], perm[:n], counts[:n_cells]

    def get_grid(self, scene_id, buf, n_nodes):
        out = np.zeros((3, n_nodes))
        _check(lib().icempm_batch_get_grid(self._h, scene_id, buf, _dptr(out), n_nodes))
        return out

    def position_map(self, scene_id):
        fb, ox, oy = C.c_int(), C.c_int(), C.c_int()
        _check(lib().icempm_batch_get_position_map(self._h, scene_id, C.byref(fb), C.byref(ox), C.byref(oy)))
        return fb.value, ox.value, oy.value

    def autotune(self, n_trials=3):
        k = 2 * len(tile_candidates())
        out = (C.c_float * k)()
        chosen = C.c_int(-1)
        _check(lib().icempm_batch_autotune(self._h, n_trials, out, k, C.byref(chosen)))
        return chosen.value, list(out)

    def set_tile_shape(self, tx, ty, margin=1):
        _check(lib().icempm_batch_set_tile_shape(self._h, tx, ty, margin))

    def tile_config(self):
        v = [C.c_int() for _ in range(4)]
        _check(lib().icempm_batch_get_tile_config(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("tx", "ty", "margin", "lane_spread"), (x.value for x in v)))

    def graph_stats(self):
        n, c = C.c_longlong(), C.c_int()
        _check(lib().icempm_batch_get_graph_stats(self._h, C.byref(n), C.byref(c)))
        return n.value, c.value

    def kernel_stats(self):
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = (C.c_longlong * len(KERNEL_CLASSES))()
        _check(lib().icempm_batch_get_kernel_stats(self._h, _dptr(ms), cnt, len(KERNEL_CLASSES)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def reset_kernel_stats(self):
        _check(lib().icempm_batch_reset_kernel_stats(self._h))

    def last_step_ms(self):
        v = C.c_float()
        _check(lib().icempm_batch_last_step_ms(self._h, C.byref(v)))
        return v.value
