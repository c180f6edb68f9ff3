"""The batched workloads at their full size.

- 64 independent ice blocks: 7.2M points, 10k-500k per scene, grids 128x128
  to 1024x512, per-scene E and cohesion.  FP64 and mixed forms.
- 8 floe fields: 24.1M points of convex floes, grids 4096x1024 to 8192x2048.
  Mixed and FP64 forms.

Each batch is autotuned and then runs two frames of 500 substeps as hipGraph
chunk replays, with a device re-sort between the frames.  Every scene is then
checked:

- error byte 0 and a finite state;
- grid mass equal to N*m to 1e-9 (N*float32(m) in the mixed form);
- the frame-boundary sort is a stable counting sort (keys non-decreasing,
  a permutation, ties in their previous order, counts equal the histogram);
- keys equal the oracle's on a seeded sample of points.

The FP64 form matches the FP64 oracle (1e-8 of each field's scale after 1000
substeps, plus 1e-12 for v and C, which stay at round-off size on floes at
rest) on the three smallest blocks, the largest block (445k points) and
the smallest floe field (1.3M points).  The mixed form stays within 1e-4
relative position error of the FP64 oracle after 1000 substeps on the same
scenes.
"""
import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

FRAME = 500
SAMPLE = 2000
FIELDS = {"x": (4, 6), "v": (6, 8), "F": (8, 12), "C": (12, 16), "Jp": (3, 4)}


@pytest.fixture(scope="module")
def blocks():
    bag = scenes.batch_scenes()
    assert len(bag) == 64 and 7_000_000 < sum(n for _, _, n in bag) < 7_500_000
    return bag


@pytest.fixture(scope="module")
def floes():
    bag = scenes.floe_fields()
    assert len(bag) == 8 and 23_500_000 < sum(n for _, _, n in bag) < 24_500_000
    return bag


def oracle_two_frames(P, h, n):
    """Sort, a frame, sort, a frame (the batch's sequence) on the CPU oracle."""
    ref = h.copy()
    O.sort(P, ref, n)
    t, step, _, err = O.substeps(P, ref, n, FRAME, 0.0, 0)
    assert err == 0
    O.sort(P, ref, n)
    _, _, _, err = O.substeps(P, ref, n, FRAME, t, step)
    assert err == 0
    return ref


_oracle_cache = {}


def oracle_ref(key, P, h, n):
    if key not in _oracle_cache:
        _oracle_cache[key] = oracle_two_frames(P, h, n)
    return _oracle_cache[key]


def block_checked(bag):
    order = np.argsort([n for _, _, n in bag])
    return [int(i) for i in order[:3]] + [int(order[-1])]


def floe_checked(bag):
    return [int(np.argmin([n for _, _, n in bag]))]


def run_two_frames(bag, precision):
    """Batch, autotune, frame 1, download, frame 2 with the re-sort first.
    Returns (batch, states before the re-sort, final states)."""
    b = im.Batch(0, precision)
    for P, h, n in bag:
        b.add_scene(P, h, n)
    b.commit()
    chosen, _ = b.autotune(2)
    assert chosen >= 0
    print("tile", b.tile_config(), flush=True)
    b.step(FRAME, im.STEP_GRAPH)
    assert b.error_code == 0
    mid = []
    for i, (P, h, n) in enumerate(bag):
        o = np.zeros_like(h)
        b.transfer_from_device(i, o)
        mid.append(o)
    b.step(FRAME, im.STEP_GRAPH | im.STEP_RESORT)
    assert b.error_code == 0
    final = []
    for i, (P, h, n) in enumerate(bag):
        o = np.zeros_like(h)
        b.transfer_from_device(i, o)
        final.append(o)
    return b, mid, final


def check_scene(b, i, P, n, mid, final, precision, rng):
    """Finite state, grid mass, and the bins of the frame-boundary sort (of
    the state `mid`) against the oracle's keys on a sample."""
    assert b.scene_points(i) == n
    assert np.all(np.isfinite(final[1:16, :n])), f"scene {i}: non-finite state"
    nn = O.n_nodes(P)
    mass = sum(b.get_grid(i, k, nn)[0].sum() for k in range(3))   # one buffer holds the last P2G
    m = P.ParticleMass if precision == im.PRECISION_FP64 else float(np.float32(P.ParticleMass))
    assert abs(mass - n * m) <= 1e-9 * n * m, f"scene {i}: grid mass {mass} vs {n * m}"
    keys, perm, counts = b.get_bins(i, n, nn)
    assert np.all(np.diff(keys) >= 0), f"scene {i}: keys not sorted"
    assert np.array_equal(np.sort(perm), np.arange(n)), f"scene {i}: not a permutation"
    same = np.diff(keys) == 0
    assert np.all(np.diff(perm)[same] > 0), f"scene {i}: ties out of their previous order"
    assert np.array_equal(counts, np.bincount(keys + P.GridHaloSize * P.GridY, minlength=nn)), f"scene {i}: counts"
    if precision == im.PRECISION_MIXED:
        fb, ox, oy = b.position_map(i)
    for k in rng.choice(n, size=min(SAMPLE, n), replace=False):
        j = perm[k]
        if precision == im.PRECISION_FP64:
            ref = O.cell_key(P, mid[4, j], mid[5, j])
        else:
            xq = int(np.rint((mid[4, j] / P.cellsize + ox) * 2.0 ** fb))
            yq = int(np.rint((mid[5, j] / P.cellsize + oy) * 2.0 ** fb))
            ref = O.cell_key_fixed(xq, yq, fb, ox, oy, P.GridY)
        assert keys[k] == ref, (i, k, j, keys[k], ref)


# Absolute floors of the comparison: floes that stay at rest keep v and C at
# round-off size (1e-14), where a bar relative to the field's own size is
# meaningless.  1e-12 m/s and 1e-12 1/s are far below any physical motion here.
FLOOR = {"v": 1e-12, "C": 1e-12}


def assert_matches_oracle(out, ref, n, what):
    for name, (a, z) in FIELDS.items():
        s = max(np.abs(ref[a:z, :n]).max(), 1e-300)
        d = np.abs(out[a:z, :n] - ref[a:z, :n]).max()
        assert d <= 1e-8 * s + FLOOR.get(name, 0.0), f"{what} {name}: max diff {d:.3e} scale {s:.3e}"
    mism = np.count_nonzero(out[0, :n].view(np.int64) != ref[0, :n].view(np.int64))
    assert mism == 0, f"{what}: {mism} utility words differ"


def rel_pos(out, ref, n):
    return np.linalg.norm(out[4:6, :n] - ref[4:6, :n]) / np.linalg.norm(ref[4:6, :n])


@pytest.mark.parametrize("precision", [im.PRECISION_FP64, im.PRECISION_MIXED], ids=["fp64", "mixed"])
def test_64_blocks_full_size(blocks, precision):
    b, mid, final = run_two_frames(blocks, precision)
    rng = np.random.default_rng(0x1CE64)
    for i, (P, h, n) in enumerate(blocks):
        check_scene(b, i, P, n, mid[i], final[i], precision, rng)
    b.close()
    for i in block_checked(blocks):
        P, h, n = blocks[i]
        ref = oracle_ref(("block", i), P, h, n)
        if precision == im.PRECISION_FP64:
            assert_matches_oracle(final[i], ref, n, f"block {i} ({n} points)")
        else:
            e = rel_pos(final[i], ref, n)
            print(f"block {i} ({n} points): mixed rel_pos after 1000 substeps {e:.3e}", flush=True)
            assert e <= 1e-4, (i, n, e)


@pytest.mark.parametrize("precision", [im.PRECISION_MIXED, im.PRECISION_FP64], ids=["mixed", "fp64"])
def test_8_floe_fields_full_size(floes, precision):
    b, mid, final = run_two_frames(floes, precision)
    rng = np.random.default_rng(0x1CE08)
    for i, (P, h, n) in enumerate(floes):
        check_scene(b, i, P, n, mid[i], final[i], precision, rng)
    b.close()
    del mid
    for i in floe_checked(floes):
        P, h, n = floes[i]
        ref = oracle_ref(("floe", i), P, h, n)
        if precision == im.PRECISION_FP64:
            assert_matches_oracle(final[i], ref, n, f"floe field {i} ({n} points)")
        else:
            e = rel_pos(final[i], ref, n)
            print(f"floe field {i} ({n} points): mixed rel_pos after 1000 substeps {e:.3e}", flush=True)
            assert e <= 1e-4, (i, n, e)
