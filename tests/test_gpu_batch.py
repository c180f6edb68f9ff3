"""Batched form: many independent scenes of uneven sizes, grids and materials
stepped by one launch per phase, against the FP64 oracle scene by scene."""
import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = {"x": 4, "y": 5, "vx": 6, "vy": 7, "F00": 8, "F01": 9, "F10": 10, "F11": 11, "C00": 12, "C11": 15,
          "Jp": 3}


def mixed_bag(seed=0x1CE0A00):
    """Six scenes: blocks of different sizes, grids, E and cohesion, one with an indenter."""
    out = []
    specs = [dict(nx=30, ny=8, grid=(64, 32), E=3e9, cohesion=5e4),
             dict(nx=60, ny=12, grid=(128, 64), E=9e9, cohesion=2e5),
             dict(nx=20, ny=20, grid=(64, 64), E=5e9, cohesion=1e5),
             dict(nx=90, ny=6, grid=(128, 32), E=7e9, cohesion=8e4),
             dict(nx=44, ny=10, grid=(96, 48), E=4e9, cohesion=1.5e5)]
    for i, s in enumerate(specs):
        out.append(scenes.block_scene(seed=seed + i, n_grains=30, **s))
    out.append(scenes.sheet_scene(seed=seed + 9, nx=100, ny=16, grid=(160, 48), h=0.05, ind_diameter=0.5,
                                  ind_depth=0.15, n_grains=30, sampler="jittered"))
    for P, h, n in out:                      # some motion so every term is exercised
        r = np.random.default_rng(n)
        h[6, :n] = 0.3 * np.sin(h[4, :n])
        h[7, :n] = -0.2 * np.cos(h[5, :n])
        h[8:12, :n] += 1e-4 * r.normal(size=(4, n))
    return out


def run_batch(bag, precision, flags, nsub=(9, 11)):
    b = im.Batch(0, precision)
    for P, h, n in bag:
        b.add_scene(P, h, n)
    b.commit()
    for k in nsub:
        b.step(k, flags)
    outs = []
    for i, (P, h, n) in enumerate(bag):
        o = np.zeros_like(h)
        b.transfer_from_device(i, o)
        outs.append(o)
    forces = [b.indenter_force(i) for i in range(len(bag))]
    return b, outs, forces


@pytest.mark.parametrize("flags", [0, im.STEP_GRAPH, im.STEP_UNFUSED, im.STEP_DIRECT_P2G])
def test_batch_fp64_matches_oracle_per_scene(flags):
    bag = mixed_bag()
    b, outs, forces = run_batch(bag, im.PRECISION_FP64, flags)
    assert b.error_code == 0
    assert b.counts()[0] == len(bag)
    for i, (P, h, n) in enumerate(bag):
        ref = h.copy()
        if not flags & im.STEP_DIRECT_P2G:
            O.sort(P, ref, n)
        _, _, fref, err = O.substeps(P, ref, n, 20, 0.0, 0)
        assert err == 0
        for name, line in FIELDS.items():
            s = np.abs(ref[line, :n]).max() + 1e-300
            d = np.abs(outs[i][line, :n] - ref[line, :n]).max()
            assert d <= 1e-9 * s, f"scene {i} {name}: {d:.3e} vs {s:.3e}"
        assert np.array_equal(outs[i][0, :n].view(np.int64), ref[0, :n].view(np.int64))
        assert np.allclose(forces[i].reshape(-1, 2).sum(0), fref.reshape(-1, 2).sum(0), rtol=1e-8, atol=1e-6)


def test_batch_mixed_accuracy():
    bag = mixed_bag(seed=0x1CE0B00)
    b, outs, _ = run_batch(bag, im.PRECISION_MIXED, im.STEP_GRAPH, nsub=(200, 300))
    assert b.error_code == 0
    for i, (P, h, n) in enumerate(bag):
        ref = h.copy()
        O.sort(P, ref, n)
        O.substeps(P, ref, n, 500, 0.0, 0)
        rel = np.linalg.norm(outs[i][4:6, :n] - ref[4:6, :n]) / np.linalg.norm(ref[4:6, :n])
        assert rel <= 1e-4, (i, rel)


def test_batch_equals_single_partitions():
    """A batch of scenes steps each scene exactly as a partition holding it alone
    would (up to the order of the FP64 atomic sums)."""
    bag = mixed_bag(seed=0x1CE0C00)[:3]
    b, outs, _ = run_batch(bag, im.PRECISION_FP64, 0)
    for i, (P, h, n) in enumerate(bag):
        p = im.Partition(0, 0, im.PRECISION_FP64)
        p.allocate(h.shape[1], P.GridXTotal)
        p.update_constants(P)
        p.transfer_points_from_soa_to_device(h, 0, n)
        p.sort_points()
        p.step(9, 0)
        p.step(11, 0)
        o = np.zeros_like(h)
        p.transfer_from_device(o, 0)
        for line in FIELDS.values():
            s = np.abs(o[line, :n]).max() + 1e-300
            assert np.abs(o[line, :n] - outs[i][line, :n]).max() <= 1e-10 * s


def test_batch_error_code_from_one_scene():
    bag = mixed_bag(seed=0x1CE0D00)[:3]
    P, h, n = bag[1]
    h[5, 7] = (P.GridY - 1.4) * P.cellsize
    b, _, _ = run_batch(bag, im.PRECISION_FP64, im.STEP_DIRECT_P2G, nsub=(1,))
    assert b.error_code == 73


def test_batch_autotune_and_resort():
    bag = mixed_bag(seed=0x1CE0E00)
    b = im.Batch(0, im.PRECISION_FP64)
    for P, h, n in bag:
        b.add_scene(P, h, n)
    b.commit()
    chosen, ms = b.autotune(2)
    assert chosen >= 0
    b.step(10, im.STEP_GRAPH)
    b.step(10, im.STEP_GRAPH | im.STEP_RESORT)
    assert b.error_code == 0
    n_sc, n_tot = b.counts()
    assert n_sc == len(bag) and n_tot == sum(n for _, _, n in bag)


def test_floe_fields_batched_mixed_and_fp64():
    """Two small floe fields (convex floes of different sizes on one grid each)
    in one batch: FP64 against the oracle, mixed within 1e-4 in position."""
    bag = [scenes.floe_field(seed=0x1CE0F00 + i, grid=(256, 128), target_points=12_000 + 4_000 * i)
           for i in range(2)]
    for P, h, n in bag:                      # motion and strain so that every term is exercised
        r = np.random.default_rng(n)
        h[6, :n] = 0.3 * np.sin(2 * h[4, :n])
        h[7, :n] = -0.2 * np.cos(3 * h[5, :n])
        h[8:12, :n] += 1e-4 * r.normal(size=(4, n))
    b, outs, _ = run_batch(bag, im.PRECISION_FP64, im.STEP_GRAPH, nsub=(50, 50))
    bm, outs_m, _ = run_batch(bag, im.PRECISION_MIXED, im.STEP_GRAPH, nsub=(50, 50))
    assert b.error_code == 0 and bm.error_code == 0
    for i, (P, h, n) in enumerate(bag):
        ref = h.copy()
        O.sort(P, ref, n)
        _, _, _, err = O.substeps(P, ref, n, 100, 0.0, 0)
        assert err == 0
        # each field group against its own magnitude
        for name, (a, z) in {"x": (4, 6), "v": (6, 8), "F": (8, 12), "C": (12, 16), "Jp": (3, 4)}.items():
            s = np.abs(ref[a:z, :n]).max() + 1e-300
            d = np.abs(outs[i][a:z, :n] - ref[a:z, :n]).max()
            assert d <= 1e-9 * s, f"floes {i} {name}: {d:.3e} vs {s:.3e}"
        rel = np.linalg.norm(outs_m[i][4:6, :n] - ref[4:6, :n]) / np.linalg.norm(ref[4:6, :n])
        assert rel <= 1e-4, (i, rel)
