"""Helpers for the tests of the C++ host layer (icempm_run driver)."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "icempm_multi_amd", "icempm_run")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def write_flat_raw(path, x, y, grains, volume):
    """Flat raw-point file of icempm_run: int64 n, float32 volume, pad, x, y, llGrainIDs."""
    n = len(x)
    with open(path, "wb") as f:
        f.write(np.array([n], dtype=np.int64).tobytes())
        f.write(np.array([volume, 0], dtype=np.float32).tobytes())
        f.write(np.asarray(x, dtype=np.float64).tobytes())
        f.write(np.asarray(y, dtype=np.float64).tobytes())
        f.write(np.asarray(grains, dtype=np.uint64).tobytes())


def read_flat_raw(path):
    b = open(path, "rb").read()
    n = int(np.frombuffer(b[:8], np.int64)[0])
    vol = np.frombuffer(b[8:12], np.float32)[0]
    x = np.frombuffer(b[16:16 + 8 * n], np.float64)
    y = np.frombuffer(b[16 + 8 * n:16 + 16 * n], np.float64)
    g = np.frombuffer(b[16 + 16 * n:16 + 24 * n], np.uint64)
    return x, y, g, vol


def read_soa_dump(path):
    """--dump-soa file: uint32 capacity, uint32 size, 16 lines of capacity doubles."""
    b = open(path, "rb").read()
    cap, size = np.frombuffer(b[:8], np.uint32)
    soa = np.frombuffer(b[8:], np.float64).reshape(16, int(cap)).copy()
    return soa, int(size)


def run(*args, cwd=None, timeout=300):
    r = subprocess.run([RUN, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise AssertionError(f"icempm_run {args} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def p1_config(tmp, **overrides):
    """The reference's p1 configuration (2d_70k points), with overrides."""
    import json
    cfg = json.load(open(os.path.join(GOLDEN, "configs", "build_p1.json")))
    cfg.update(overrides)
    path = os.path.join(tmp, "p1.json")
    with open(path, "w") as f:
        json.dump(cfg, f)
    return path
