"""ctypes loader of the FP64 CPU oracle (oracle/liboracle_mpm.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, and only as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from icempm_multi_amd.params import SimParams, N_PTS_ARRAYS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle_mpm.so")

_lib = None

D = C.POINTER(C.c_double)
P_ = C.POINTER(SimParams)


def _ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    L = C.CDLL(ORACLE_SO)
    sig = {
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_params_reset": (None, [P_]),
        "orc_params_compute_lame": (None, [P_]),
        "orc_params_camclay": (None, [P_]),
        "orc_params_helper": (None, [P_]),
        "orc_weights": (None, [C.c_double, D]),
        "orc_kirchhoff_stress": (None, [P_, D, D]),
        "orc_water_stress": (None, [P_, C.c_double, D]),
        "orc_compute_pq": (None, [P_, D, D, D, D]),
        "orc_grain_params": (None, [P_, C.c_int, D]),
        "orc_failure_check": (C.c_int, [P_, C.c_int, C.c_double, C.c_double]),
        "orc_polar2x2": (None, [D, D, D]),
        "orc_svd2x2": (None, [D, D, D, D]),
        "orc_point_plasticity": (None, [P_, D, D, C.POINTER(C.c_int64), D, D]),
        "orc_p2g": (C.c_int, [P_, D, C.c_size_t, C.c_int, D, C.c_size_t, C.POINTER(C.c_uint8)]),
        "orc_update_nodes": (None, [P_, D, C.c_size_t, C.c_double, C.c_double, C.c_double, D]),
        "orc_g2p": (C.c_int, [P_, D, C.c_size_t, C.c_int, D, C.c_size_t, C.c_int, C.POINTER(C.c_uint8)]),
        "orc_substeps": (C.c_int, [P_, D, C.c_size_t, C.c_int, C.c_int, D, C.POINTER(C.c_int), D,
                                   C.POINTER(C.c_uint8)]),
        "orc_cell_key": (C.c_int, [P_, C.c_double, C.c_double]),
        "orc_sort": (C.c_int, [P_, D, C.c_size_t, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "orc_cell_key_fixed": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_simparams_offsets": (C.c_int, [C.POINTER(C.c_size_t), C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def reset_params():
    P = SimParams()
    lib().orc_params_reset(C.byref(P))
    return P


def derive(P):
    """ComputeCamClayParams2 + ComputeHelperVariables, as ParseFile ends."""
    lib().orc_params_camclay(C.byref(P))
    lib().orc_params_helper(C.byref(P))
    return P


def n_nodes(P):
    return P.GridY * (P.GridXTotal + 2 * P.GridHaloSize)


def weights(fx):
    w = np.zeros(3)
    lib().orc_weights(fx, _ptr(w))
    return w


def kirchhoff(P, F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
    t = np.zeros(4)
    lib().orc_kirchhoff_stress(C.byref(P), _ptr(F), _ptr(t))
    return t


def water(P, J):
    t = np.zeros(4)
    lib().orc_water_stress(C.byref(P), J, _ptr(t))
    return t


def compute_pq(P, F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
    je, p, q = C.c_double(), C.c_double(), C.c_double()
    lib().orc_compute_pq(C.byref(P), _ptr(F), C.byref(je), C.byref(p), C.byref(q))
    return je.value, p.value, q.value


def grain_params(P, grain):
    o = np.zeros(6)
    lib().orc_grain_params(C.byref(P), grain, _ptr(o))
    return dict(zip(["pmin", "pmax", "qmax", "beta", "mSq", "pmin2"], o))


def failure_check(P, grain, p, q):
    return lib().orc_failure_check(C.byref(P), grain, p, q)


def polar(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
    R, S = np.zeros(4), np.zeros(4)
    lib().orc_polar2x2(_ptr(F), _ptr(R), _ptr(S))
    return R.reshape(2, 2), S.reshape(2, 2)


def svd(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(4)
    U, S, V = np.zeros(4), np.zeros(2), np.zeros(4)
    lib().orc_svd2x2(_ptr(F), _ptr(U), _ptr(S), _ptr(V))
    return U.reshape(2, 2), S, V.reshape(2, 2)


def point_plasticity(P, F, Jp_inv, utility):
    F = np.array(F, dtype=np.float64).reshape(4)
    jp = C.c_double(Jp_inv)
    ud = C.c_int64(utility)
    p, q = C.c_double(), C.c_double()
    lib().orc_point_plasticity(C.byref(P), _ptr(F), C.byref(jp), C.byref(ud), C.byref(p), C.byref(q))
    return F.reshape(2, 2), jp.value, ud.value, p.value, q.value


def p2g(P, soa, n):
    pitch = soa.shape[1]
    nn = n_nodes(P)
    grid = np.zeros((3, nn))
    err = C.c_uint8(0)
    lib().orc_p2g(C.byref(P), _ptr(soa), pitch, n, _ptr(grid), nn, C.byref(err))
    return grid, err.value


def update_nodes(P, grid, t, ind_x, ind_y):
    force = np.zeros(2 * P.n_indenter_subdivisions)
    lib().orc_update_nodes(C.byref(P), _ptr(grid), grid.shape[1], t, ind_x, ind_y, _ptr(force))
    return force


def g2p(P, soa, n, grid, recordPQ=0):
    err = C.c_uint8(0)
    lib().orc_g2p(C.byref(P), _ptr(soa), soa.shape[1], n, _ptr(grid), grid.shape[1], recordPQ, C.byref(err))
    return err.value


def substeps(P, soa, n, nsub, t, step):
    """Runs nsub substeps in place; returns (t, step, indenter_force, err)."""
    tt = C.c_double(t)
    st = C.c_int(step)
    force = np.zeros(2 * P.n_indenter_subdivisions)
    err = C.c_uint8(0)
    lib().orc_substeps(C.byref(P), _ptr(soa), soa.shape[1], n, nsub, C.byref(tt), C.byref(st), _ptr(force),
                       C.byref(err))
    return tt.value, st.value, force, err.value


def cell_key(P, x, y):
    return lib().orc_cell_key(C.byref(P), x, y)


def sort(P, soa, n):
    keys = np.zeros(max(n, 1), dtype=np.int32)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    m = lib().orc_sort(C.byref(P), _ptr(soa), soa.shape[1], n, _ptr(keys, C.c_int32), _ptr(perm, C.c_int32))
    return m, keys[:m], perm[:m]


def cell_key_fixed(xq, yq, frac_bits, origin_x, origin_y, GridY):
    return lib().orc_cell_key_fixed(int(xq), int(yq), frac_bits, origin_x, origin_y, GridY)


def simparams_offsets():
    out = (C.c_size_t * 128)()
    m = lib().orc_simparams_offsets(out, 128)
    return list(out[:m])


def empty_soa(capacity):
    return np.zeros((N_PTS_ARRAYS, capacity), dtype=np.float64)
