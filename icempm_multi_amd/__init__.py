"""MI355X-native 2D MPM ice stepper (drop-in for the simulation step of
igorg520b/iceMPM_multi).

The stepping runs in hand-written HIP kernels for gfx950 inside
``libicempm.so`` (built from ``icempm_multi_amd/csrc``); this package is its
ctypes binding plus seeded scene generators.  Nothing here computes the step
on the CPU: without the built library every call raises.
"""
from .params import (SimParams, FIELD_NAMES, N_PTS_ARRAYS, PRECISION_FP64, PRECISION_MIXED, STATUS_CRUSHED,
                     STATUS_DISABLED, STATUS_LIQUID, STEP_DIRECT_P2G, STEP_GRAPH, STEP_RESORT, STEP_TIMINGS,
                     STEP_SAMPLED, STEP_SPAN, STEP_UNFUSED, OK, ERR_ARGUMENT, ERR_CAPACITY, ERR_DEVICE, ERR_HIP, ERR_NOMEM,
                     ERR_STATE)
from .capi import (Batch, IcempmError, Partition, EXPORTED_SYMBOLS, LIB_PATH, lib, params_derive,
                   params_parse_json, params_reset, prepare_raw_points)

__all__ = [
    "SimParams", "FIELD_NAMES", "N_PTS_ARRAYS", "PRECISION_FP64", "PRECISION_MIXED", "STATUS_CRUSHED",
    "STATUS_DISABLED", "STATUS_LIQUID", "STEP_DIRECT_P2G", "STEP_GRAPH", "STEP_RESORT", "STEP_TIMINGS",
    "STEP_SAMPLED", "STEP_SPAN", "STEP_UNFUSED", "Batch", "IcempmError", "Partition", "EXPORTED_SYMBOLS", "LIB_PATH", "lib", "params_derive",
    "params_parse_json", "params_reset", "prepare_raw_points", "OK", "ERR_ARGUMENT", "ERR_CAPACITY", "ERR_DEVICE",
    "ERR_HIP", "ERR_NOMEM", "ERR_STATE",
]
