"""CPU checks of the C-ABI library: it loads without a GPU, exports every
function include/icempm.h declares, and its host-only logic (parameters, JSON,
LoadRawPoints preparation) follows the reference."""
import json
import math
import os
import re

import numpy as np
import pytest

import icempm_multi_amd as im
from icempm_multi_amd import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def header_functions():
    text = open(os.path.join(ROOT, "include", "icempm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(icempm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = im.lib()
    names = header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTED_SYMBOLS)


def test_python_constants_match_header_defines():
    # the Python names of the status codes, step flags, precisions and point
    # status bits carry the values include/icempm.h defines
    text = open(os.path.join(ROOT, "include", "icempm.h")).read()
    defs = {m.group(1): int(m.group(2), 0)
            for m in re.finditer(r"#define\s+ICEMPM_([A-Z0-9_]+)\s+\(?(-?(?:0x[0-9A-Fa-f]+|\d+))\)?", text)}
    checked = 0
    for name, value in defs.items():
        if hasattr(im, name):
            assert getattr(im, name) == value, name
            checked += 1
    for name in ("ERR_ARGUMENT", "ERR_STATE", "ERR_DEVICE", "STEP_GRAPH", "STEP_SAMPLED", "PRECISION_MIXED"):
        assert name in defs and hasattr(im, name), name
    assert checked >= 14


def test_struct_size_matches_header_mirror():
    import ctypes
    from _oracle import simparams_offsets
    assert ctypes.sizeof(im.SimParams) == simparams_offsets()[-1]


def test_reset_matches_oracle_reset():
    import _oracle as O
    a, b = im.params_reset(), O.reset_params()
    assert bytes(a) == bytes(b)


@pytest.mark.parametrize("cfg", ["build_p1.json", "build_p2.json", "input_p1.json", "input_p2.json"])
def test_parse_json_reference_configs(cfg):
    text = open(os.path.join(GOLDEN, "configs", cfg)).read()
    d = json.loads(text)
    P, raw = im.params_parse_json(text)
    assert raw == d["InputRawPoints"]
    assert P.GridXTotal == d["GridX"] and P.GridY == d["GridY"] and P.GridHaloSize == d["GridHaloSize"]
    assert P.InitialTimeStep == d["InitialTimeStep"] and P.YoungsModulus == d["YoungsModulus"]
    assert P.DP_tan_phi == math.tan(d["DP_phi"] * math.pi / 180)
    assert P.DP_threshold_p == d["DP_threshold_p"]
    assert P.GrainVariability == d.get("GrainVariability", 0.5)
    assert P.cellsize == d["GridXDimension"] / d["GridX"]
    assert P.UpdateEveryNthStep == int(round(1 / (200 * d["InitialTimeStep"])))
    assert math.isclose(P.kappa, d["YoungsModulus"] / (3 * (1 - 2 * d["PoissonsRatio"])), rel_tol=1e-14)
    assert P.tpb_P2G == 256 and P.nPartitions == d["nPartitions"]


def test_parse_json_errors():
    with pytest.raises(im.IcempmError):
        im.params_parse_json('{"GridX": 10}')            # InputRawPoints missing
    with pytest.raises(im.IcempmError):
        im.params_parse_json("not json")


def test_prepare_raw_points_like_load_raw_points():
    d = np.load(os.path.join(GOLDEN, "ref_2d_70k.npz"))
    P, _ = im.params_parse_json(open(os.path.join(GOLDEN, "configs", "build_p1.json")).read())
    x, y, g = d["x"], d["y"], d["llGrainIDs"]
    h, n = scenes.load_raw_points(P, x, y, g, float(d["volume"][0]))
    assert n == len(x) and h.shape[1] == int(n * 1.2)
    hh = P.cellsize
    off = np.array([(P.GridXTotal * hh - (x.max() - x.min())) / 2, 2 * hh])
    # same multiset of (x, y, grain), offset, ordered by the cell key, ties in file order
    keys = (((x + off[0]) / hh + 0.5).astype(int) * P.GridY + ((y + off[1]) / hh + 0.5).astype(int))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(h[4, :n], (x + off[0])[order]) and np.array_equal(h[5, :n], (y + off[1])[order])
    assert np.array_equal(h[0, :n].view(np.uint64), g[order])
    assert np.all(h[8, :n] == 1) and np.all(h[11, :n] == 1) and np.all(h[9:11, :n] == 0) and np.all(h[3, :n] == 1)
    assert math.isclose(P.ParticleVolume, 2.5 / n, rel_tol=1e-15)
    assert math.isclose(P.ParticleMass, 980 * 2.5 / n, rel_tol=1e-15)
    r = 0.324 / 2
    ht = r - 0.101
    assert math.isclose(P.indenter_x, math.floor((off[0] - math.sqrt(r * r - ht * ht)) / hh) * hh, rel_tol=1e-15)
    assert math.isclose(P.indenter_y, y.max() + off[1] + ht, rel_tol=1e-15)


def test_scene_generators_are_seeded():
    P1, a, n1 = scenes.block_scene(nx=20, ny=8, grid=(64, 32), n_grains=10)
    P2, b, n2 = scenes.block_scene(nx=20, ny=8, grid=(64, 32), n_grains=10)
    assert n1 == n2 == 640 and np.array_equal(a, b)
    assert math.isclose(P1.DP_threshold_p, 1e5 / math.tan(math.radians(30)))
    u = scenes.uniform(42, 1000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05


def test_no_device_calls_fail_cleanly():
    # without a GPU the device entry points report an error, never crash
    try:
        p = im.Partition(0, 0, im.PRECISION_FP64)
    except im.IcempmError as e:
        assert e.code < 0
    else:
        p.close()


def test_tile_candidates():
    """The autotuner's shapes: within the kernels' limits, the earlier powers of
    two first (their results keep their indices), then the widths between
    them that let a scene's tile count come close to whole rounds."""
    c = capi.tile_candidates()
    assert len(c) == capi.lib().icempm_tile_candidates(None, None, None, 0) >= 14
    assert c[12] == (8, 64, 1) and c[11] == (16, 32, 1)
    assert len(set(c)) == len(c)
    assert all(1 <= tx <= 32 and ty in (8, 16, 32, 64) and ls in (0, 1, 2) for tx, ty, ls in c)
    assert {(6, 64, 1), (7, 64, 1), (12, 32, 1)} <= set(c)
