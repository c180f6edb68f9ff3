"""FP64 HIP path vs the FP64 CPU oracle, through the C-ABI.

Differences come only from the order of the FP64 atomic sums in P2G (the
kernels are built without FP contraction), so fields must agree to ~1e-10 of
their scale after tens of substeps; binning (keys, bin counts, order) must be
bit-exact.
"""
import os

import numpy as np
import pytest

import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes
from icempm_multi_amd.params import STATUS_CRUSHED, STATUS_DISABLED, STATUS_LIQUID

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = {"x": 4, "y": 5, "vx": 6, "vy": 7, "F00": 8, "F01": 9, "F10": 10, "F11": 11,
          "C00": 12, "C01": 13, "C10": 14, "C11": 15, "Jp": 3}


def make_partition(P, hssoa, n, precision=im.PRECISION_FP64):
    part = im.Partition(0, 0, precision)
    part.allocate(hssoa.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(hssoa, 0, n)
    return part


def download(part, like):
    out = np.zeros_like(like)
    part.transfer_from_device(out, 0)
    return out


def assert_close(out, ref, n, rtol=1e-10, fields=FIELDS, what=""):
    for name, line in fields.items():
        s = np.abs(ref[line, :n]).max() + 1e-300
        d = np.abs(out[line, :n] - ref[line, :n]).max()
        assert d <= rtol * s, f"{what} {name}: max diff {d:.3e} scale {s:.3e}"
    ud_o = out[0, :n].view(np.int64)
    ud_r = ref[0, :n].view(np.int64)
    assert np.array_equal(ud_o, ud_r), f"{what} utility_data differs at {np.nonzero(ud_o != ud_r)[0][:10]}"


def small_block(seed=0x1CE0000 + 7, **kw):
    kw.setdefault("nx", 48)
    kw.setdefault("ny", 12)
    kw.setdefault("grid", (96, 48))
    kw.setdefault("n_grains", 60)
    return scenes.block_scene(seed=seed, **kw)


def with_motion(P, hssoa, n, seed=3):
    """Adds a velocity field and small strains so every term of the step is exercised."""
    r = np.random.default_rng(seed)
    hssoa[6, :n] = 0.5 + 0.2 * np.sin(hssoa[4, :n])
    hssoa[7, :n] = -0.3 * np.cos(hssoa[5, :n])
    hssoa[8:12, :n] += 2e-4 * r.normal(size=(4, n))
    hssoa[12:16, :n] = 0.5 * r.normal(size=(4, n))
    return hssoa


def test_phases_unbinned_match_oracle():
    P, h, n = small_block()
    with_motion(P, h, n)
    part = make_partition(P, h, n)
    nn = O.n_nodes(P)
    part.reset_grid()
    part.p2g()
    g_gpu = part.get_grid(0, nn)
    g_ref, err = O.p2g(P, h, n)
    assert err == 0
    sc = np.abs(g_ref).max(axis=1, keepdims=True)
    assert np.all(np.abs(g_gpu - g_ref) <= 1e-12 * sc)
    t, ix, iy = 1e-5, -1.0, -1.0
    part.update_nodes(t, ix, iy)
    O.update_nodes(P, g_ref, t, ix, iy)
    g_gpu = part.get_grid(0, nn)
    sc = np.abs(g_ref).max(axis=1, keepdims=True)
    assert np.all(np.abs(g_gpu - g_ref) <= 1e-11 * sc)
    part.g2p(1)
    out = download(part, h)
    ref = h.copy()
    O.g2p(P, ref, n, g_ref, 1)
    assert_close(out, ref, n, rtol=1e-10, fields={**FIELDS, "P": 1, "Q": 2}, what="g2p")


def test_phases_binned_match_oracle():
    P, h, n = small_block(seed=0x1CE0100)
    with_motion(P, h, n, seed=5)
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    O.sort(P, ref, n)
    nn = O.n_nodes(P)
    part.reset_grid()
    part.p2g()
    g_gpu = part.get_grid(0, nn)
    g_ref, _ = O.p2g(P, ref, n)
    sc = np.abs(g_ref).max(axis=1, keepdims=True)
    assert np.all(np.abs(g_gpu - g_ref) <= 1e-12 * sc)
    part.update_nodes(2e-5, -1.0, -1.0)
    O.update_nodes(P, g_ref, 2e-5, -1.0, -1.0)
    part.g2p(0)
    O.g2p(P, ref, n, g_ref, 0)
    assert_close(download(part, h), ref, n, rtol=1e-10, what="binned g2p")


@pytest.mark.parametrize("flags,sort", [(im.STEP_DIRECT_P2G, False), (im.STEP_UNFUSED, True), (0, True),
                                        (im.STEP_GRAPH, True), (im.STEP_TIMINGS, True),
                                        (im.STEP_GRAPH | im.STEP_TIMINGS, True),
                                        (im.STEP_GRAPH | im.STEP_TIMINGS | im.STEP_SAMPLED, True),
                                        (im.STEP_TIMINGS | im.STEP_SAMPLED, True)])
def test_substeps_match_oracle(flags, sort):
    P, h, n = small_block(seed=0x1CE0200)
    with_motion(P, h, n, seed=8)
    ref = h.copy()
    part = make_partition(P, h, n)
    if sort:
        part.sort_points()
        O.sort(P, ref, n)
    part.step(7, flags)
    part.step(13, flags)
    t, step, _, err = O.substeps(P, ref, n, 20, 0.0, 0)
    assert err == 0 and part.error_code == 0
    assert_close(download(part, h), ref, n, rtol=1e-9, what=f"flags={flags}")
    tg, sg = part.clock
    assert sg == 20 and tg == t


def liquid_and_crushed(h, n, seed):
    """Marks the lowest 30% of the points liquid and a fifth of the rest crushed,
    with Jp_inv on both sides of 1 so that both branches of the return map run."""
    r = np.random.default_rng(seed)
    ud = h[0, :n].view(np.int64)
    liq = h[5, :n] < np.quantile(h[5, :n], 0.3)
    cr = ~liq & (r.random(n) < 0.2)
    ud[liq] |= STATUS_LIQUID
    ud[cr] |= STATUS_CRUSHED
    h[3, :n] = np.where(cr, 0.9 + 0.2 * r.random(n), 1.0)
    return liq, cr


@pytest.mark.parametrize("flags", [0, im.STEP_GRAPH, im.STEP_UNFUSED, im.STEP_DIRECT_P2G])
def test_liquid_and_crushed_points_match_oracle(flags):
    """Liquid points (water mass and stress, Jp_inv *= det(I + dt C);
    kernels.cuh:37-44, 61-63, 380) and points crushed from the start (the
    Drucker-Prager return map every substep) against the oracle."""
    P, h, n = small_block(seed=0x1CE0202)
    with_motion(P, h, n, seed=9)
    liq, cr = liquid_and_crushed(h, n, seed=5)
    assert liq.sum() > 500 and cr.sum() > 200
    ref = h.copy()
    part = make_partition(P, h, n)
    if not flags & im.STEP_DIRECT_P2G:
        part.sort_points()
        O.sort(P, ref, n)
    part.step(7, flags)
    part.step(13, flags)
    t, step, _, err = O.substeps(P, ref, n, 20, 0.0, 0)
    assert err == 0 and part.error_code == 0
    out = download(part, h)
    assert_close(out, ref, n, rtol=1e-9, what=f"liquid+crushed flags={flags}")
    assert np.array_equal(out[0, :n].view(np.int64), ref[0, :n].view(np.int64))
    liq_now = (ref[0, :n].view(np.int64) & STATUS_LIQUID) != 0
    assert np.abs(ref[3, :n][liq_now] - 1.0).max() > 1e-9      # Jp_inv of liquid points moved


@pytest.mark.parametrize("flags", [0, im.STEP_DIRECT_P2G])
def test_failure_surface_every_grain_remainder_matches_oracle(flags):
    """The failure check (kernels.cuh:560-619) reads its grain terms from a
    per-scene table indexed by (grain + 4) % 11 and (grain + 3) % 30 of the
    grain id as a signed 16-bit value.  Points with grain ids over the whole
    16-bit range (negative as shorts above 32767, so negative remainders) and
    strains log-uniform from 3e-7 to 1e-3, which put them on both sides of the
    surface (30% crush in the oracle, of both signs of grain), crush exactly
    where the oracle's do."""
    P, h, n = small_block(seed=0x1CE0203)
    with_motion(P, h, n, seed=10)
    r = np.random.default_rng(12)
    ud = h[0, :n].view(np.int64)
    ud[:] = (ud & ~np.int64(0xffff)) | r.integers(0, 65536, n)
    s = 10.0 ** r.uniform(-6.5, -3.0, n)
    h[8:12, :n] = np.array([1.0, 0, 0, 1.0])[:, None] + s * r.normal(size=(4, n))
    ref = h.copy()
    part = make_partition(P, h, n)
    if not flags & im.STEP_DIRECT_P2G:
        part.sort_points()
        O.sort(P, ref, n)
    part.step(2, flags)
    t, step, _, err = O.substeps(P, ref, n, 2, 0.0, 0)
    assert err == 0 and part.error_code == 0
    out = download(part, h)
    u_ref = ref[0, :n].view(np.int64)
    assert np.array_equal(out[0, :n].view(np.int64), u_ref)
    crushed = (u_ref & STATUS_CRUSHED) != 0
    grain = (u_ref & 0xffff).astype(np.int16)
    assert 0.1 * n < crushed.sum() < 0.9 * n, crushed.sum()
    for sel in (grain < 0, grain >= 0):
        assert crushed[sel].any() and (~crushed[sel]).any()
    assert_close(out, ref, n, rtol=1e-9, what=f"failure surface flags={flags}")


@pytest.mark.parametrize("side", ["left", "right"])
def test_walls_and_vmax_clamp_match_oracle(side):
    """The grid update's side-wall rules (kernels.cuh:290-294) with the block
    pushed against the left or right wall, and the vmax clamp (261) on a
    cluster moving at 3e4 m/s (vmax = 0.5 h / dt = 12,500 m/s here)."""
    P, h, n = small_block(seed=0x1CE0203)
    x = h[4, :n]
    if side == "left":
        x -= x.min() - 0.55          # first column of points within the wall nodes
        h[6, :n] = -0.5
    else:
        x += (P.GridXTotal * P.cellsize - 0.55) - x.max()
        h[6, :n] = 0.5
    c = np.array([x.mean(), h[5, :n].mean()])
    fast = np.hypot(x - c[0], h[5, :n] - c[1]) < 1.0
    assert fast.sum() > 20
    h[6, :n][fast] = 3e4
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    O.sort(P, ref, n)
    part.step(3, im.STEP_GRAPH)
    t, step, _, err = O.substeps(P, ref, n, 3, 0.0, 0)
    assert err == 0 and part.error_code == 0
    out = download(part, h)
    assert_close(out, ref, n, rtol=1e-8, what=f"walls/vmax {side}")
    # the clamp acted: the fast points' speed is bounded by vmax after G2P
    v = np.hypot(ref[6, :n], ref[7, :n])
    assert v.max() <= P.vmax * (1 + 1e-12) and v.max() > 0.5 * P.vmax


@pytest.mark.parametrize("graph", [0, im.STEP_GRAPH])
def test_sampled_timings_count_launches(graph):
    """Events around the launches of substeps 0, 32, 64 of a 70-substep call:
    the P2G of substep 0 and three fused launches (the grid update folded in,
    one launch per substep); a replay of the timed graph adds as many again."""
    P, h, n = small_block(seed=0x1CE0201)
    part = make_partition(P, h, n)
    part.sort_points()
    part.reset_indenter_force_accumulator()
    flags = graph | im.STEP_TIMINGS | im.STEP_SAMPLED
    part.step(70, flags)
    st = part.kernel_stats()
    assert st["p2g"][1] == 1 and st["fused_g2p_p2g"][1] == 3
    assert st["update"][1] == 0 and st["g2p"][1] == 0
    assert all(st[k][0] > 0 for k in ("p2g", "fused_g2p_p2g"))
    part.step(70, flags)
    st = part.kernel_stats()
    assert st["p2g"][1] == 2 and st["fused_g2p_p2g"][1] == 6
    part.step(70, im.STEP_GRAPH)   # untimed: no more events
    assert part.kernel_stats()["fused_g2p_p2g"][1] == 6
    assert part.error_code == 0


@pytest.mark.parametrize("graph", [0, im.STEP_GRAPH])
def test_span_timings_count_launches(graph):
    """One event pair around each run of substep launches (a graph chunk's,
    or the call's with direct launches), counted as that many launches: a
    70-substep call is the head P2G, 69 chain launches and the last substep.
    The mean of the back-to-back launches is of the size a per-launch event
    pair measures."""
    P, h, n = small_block(seed=0x1CE0202)
    part = make_partition(P, h, n)
    part.sort_points()
    part.reset_indenter_force_accumulator()
    part.step(70, graph | im.STEP_TIMINGS | im.STEP_SPAN)
    st = part.kernel_stats()
    assert st["p2g"][1] == 1 and st["fused_g2p_p2g"][1] == 69 and st["g2p"][1] == 1
    assert st["update"][1] == 0
    span_mean = st["fused_g2p_p2g"][0] / 69
    part.step(70, graph | im.STEP_TIMINGS | im.STEP_SPAN)
    st = part.kernel_stats()
    assert st["p2g"][1] == 2 and st["fused_g2p_p2g"][1] == 138 and st["g2p"][1] == 2
    part.reset_indenter_force_accumulator()
    part.step(70, graph | im.STEP_TIMINGS)
    st = part.kernel_stats()
    assert st["fused_g2p_p2g"][1] == 69
    each_mean = st["fused_g2p_p2g"][0] / 69
    assert 0.2 * each_mean < span_mean < 2.0 * each_mean, (span_mean, each_mean)
    part.step(70, graph)   # untimed: no more events
    assert part.kernel_stats()["fused_g2p_p2g"][1] == 69
    assert part.error_code == 0


@pytest.mark.parametrize("dynamic", [0, 1])
@pytest.mark.parametrize("lane_map", [0, 1, 2])
@pytest.mark.parametrize("tile", [(4, 32), (8, 16), (16, 16)])
def test_lane_maps_and_tiles_match_oracle(lane_map, tile, dynamic):
    """Every tile shape, lane-to-point map and tile distribution of the binned
    kernels, with a point count that leaves ragged last groups."""
    P, h, n = small_block(seed=0x1CE0210, nx=47, ny=13)
    with_motion(P, h, n, seed=9)
    ref = h.copy()
    part = make_partition(P, h, n)
    part.set_tile_shape(*tile)
    part.set_lane_map(lane_map)
    part.set_tile_schedule(dynamic)
    part.sort_points()
    O.sort(P, ref, n)
    part.step(15, im.STEP_GRAPH)
    _, _, _, err = O.substeps(P, ref, n, 15, 0.0, 0)
    assert err == 0 and part.error_code == 0
    assert_close(download(part, h), ref, n, rtol=1e-9, what=f"tile={tile} lane_map={lane_map}")


def test_frame_records_pq_and_clock():
    P, h, n = small_block(seed=0x1CE0300, dt=2e-4, E=1e8)     # UpdateEveryNthStep = 25, CFL ~0.1
    assert P.UpdateEveryNthStep == 25
    with_motion(P, h, n, seed=9)
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    O.sort(P, ref, n)
    part.step(30, im.STEP_GRAPH)
    t, step, _, _ = O.substeps(P, ref, n, 30, 0.0, 0)
    out = download(part, h)
    assert_close(out, ref, n, rtol=1e-9, fields={**FIELDS, "P": 1, "Q": 2}, what="pq")
    assert np.abs(out[2, :n]).max() > 0


def test_sort_bit_exact_with_disabled_and_ties():
    P, h, n = small_block(seed=0x1CE0400)
    r = np.random.default_rng(2)
    # shuffle, duplicate some positions (ties), disable a few, exact half-cell positions
    perm = r.permutation(n)
    h[:, :n] = h[:, perm]
    h[4, 100:140] = h[4, 0:40]
    h[5, 100:140] = h[5, 0:40]
    h[4, 200:220] = (np.floor(h[4, 200:220] / P.cellsize) + 0.5) * P.cellsize
    ud = h[0, :n].view(np.int64)
    ud[r.uniform(size=n) < 0.03] |= STATUS_DISABLED
    h[8, :n] = np.arange(n)          # payload to follow the permutation
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    m, keys, order = O.sort(P, ref, n)
    assert part.nPts_partition == m
    ncells = (P.GridXTotal + 2 * P.GridHaloSize) * P.GridY
    k_gpu, p_gpu, counts = part.get_bins(m, ncells)
    assert np.array_equal(k_gpu, keys)
    assert np.array_equal(p_gpu, order)
    bins = keys + P.GridHaloSize * P.GridY
    assert np.array_equal(counts, np.bincount(bins, minlength=ncells))
    out = download(part, h)
    assert np.array_equal(out[:, :m], ref[:, :m])


def test_sort_bit_exact_after_evolution():
    P, h, n = small_block(seed=0x1CE0500)
    with_motion(P, h, n, seed=4)
    h[6, :n] *= 40.0                 # move points across cells
    part = make_partition(P, h, n)
    part.step(40, im.STEP_DIRECT_P2G)
    state = download(part, h)
    ref = state.copy()
    part.sort_points()
    m, keys, order = O.sort(P, ref, n)
    k_gpu, p_gpu, _ = part.get_bins(m, 0)
    assert np.array_equal(k_gpu, keys) and np.array_equal(p_gpu, order)
    assert np.array_equal(download(part, h)[:, :m], ref[:, :m])


@pytest.mark.parametrize("where,code", [("left", 70), ("bottom", 71), ("right", 72), ("top", 73)])
def test_error_codes(where, code):
    P, h, n = small_block(seed=0x1CE0600)
    halo, hh = P.GridHaloSize, P.cellsize
    pos = {"left": (-(halo + 1.6) * hh, 2.0), "bottom": (5.0, -1.6 * hh),
           "right": ((P.GridXTotal + halo - 0.4) * hh, 2.0), "top": (5.0, (P.GridY - 1.4) * hh)}[where]
    h[4, 5], h[5, 5] = pos
    part = make_partition(P, h, n)
    part.step(1, im.STEP_DIRECT_P2G)
    assert part.error_code == code
    ref = h.copy()
    _, _, _, err = O.substeps(P, ref, n, 1, 0.0, 0)
    assert err == code


def test_indenter_force_matches_oracle():
    P, h, n = scenes.sheet_scene(nx=120, ny=20, grid=(192, 64), h=0.05, ind_diameter=0.6, ind_depth=0.2,
                                 n_grains=40, sampler="jittered")
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    O.sort(P, ref, n)
    # start with the indenter already touching the block
    part.reset_indenter_force_accumulator()
    part.step(60, 0)
    f_gpu = part.indenter_force(P.n_indenter_subdivisions)
    _, _, f_ref, err = O.substeps(P, ref, n, 60, 0.0, 0)
    assert err == 0
    tot_g, tot_r = f_gpu.reshape(-1, 2).sum(0), f_ref.reshape(-1, 2).sum(0)
    assert np.abs(tot_r).max() > 0
    assert np.allclose(tot_g, tot_r, rtol=1e-8)
    # per-bin: equal except where FP32 atan2f puts a node in the neighbouring bin
    assert np.abs(f_gpu - f_ref).sum() <= 1e-6 * np.abs(f_ref).sum() + 2 * np.abs(f_ref).max()
    assert_close(download(part, h), ref, n, rtol=1e-8, what="indenter")


def test_reference_input_2d_70k_p1():
    d = np.load(os.path.join(GOLDEN, "ref_2d_70k.npz"))
    import json
    text = open(os.path.join(GOLDEN, "configs", "build_p1.json")).read()
    P, raw = im.params_parse_json(text)
    assert raw == "geometries_2d/2d_70k.h5"
    h, n = scenes.load_raw_points(P, d["x"], d["y"], d["llGrainIDs"], float(d["volume"][0]))
    assert n == 70534
    ref = h.copy()
    part = make_partition(P, h, n)
    part.sort_points()
    O.sort(P, ref, n)
    assert np.array_equal(download(part, h)[:, :n], ref[:, :n])     # already sorted by LoadRawPoints
    part.step(25, im.STEP_GRAPH)
    O.substeps(P, ref, n, 25, 0.0, 0)
    out = download(part, h)
    assert part.error_code == 0
    assert_close(out, ref, n, rtol=1e-8, what="2d_70k/p1")


def test_autotune_keeps_state():
    P, h, n = small_block(seed=0x1CE0700)
    with_motion(P, h, n, seed=6)
    a = make_partition(P, h, n)
    a.sort_points()
    chosen, ms = a.autotune(2)
    assert chosen >= 0 and all(m > 0 for m in ms[:8] if m >= 0)
    assert a.clock == (0.0, 0)
    a.step(10, 0)
    b = make_partition(P, h, n)
    b.sort_points()
    b.step(10, 0)
    assert_close(download(a, h), download(b, h), n, rtol=1e-10, what="autotune")


def test_mass_conservation_sheet_fused():
    P, h, n = scenes.sheet_scene(nx=400, ny=40, grid=(512, 128), h=0.05, n_grains=100, sampler="jittered")
    part = make_partition(P, h, n)
    part.sort_points()
    part.step(50, im.STEP_GRAPH)
    nn = O.n_nodes(P)
    part.reset_grid()
    part.p2g()
    g = part.get_grid(0, nn)
    assert abs(g[0].sum() - n * P.ParticleMass) <= 1e-12 * n * P.ParticleMass
    assert part.error_code == 0


def test_point_capacity_limit_is_enforced():
    """Point lines are addressed with 32-bit byte offsets: a point space of
    2^29 or more is refused before anything is allocated or launched."""
    P, h, n = small_block(seed=0x1CE0800)
    part = im.Partition(0, 0, im.PRECISION_FP64)
    with pytest.raises(im.IcempmError):
        part.allocate(1 << 29, P.GridXTotal)
        part.update_constants(P)
        part.transfer_points_from_soa_to_device(h, 0, n)
    part.close()


def test_graph_chunks_survive_resort_and_stay_bounded():
    """Frames replayed as chunk graphs (1-64 substeps each) with a device
    re-sort at every frame boundary: the graphs captured for the first frames
    are replayed by the later ones, the cache stays bounded on long calls, and
    the state matches direct launches and the oracle (which re-sorts at the
    same boundaries)."""
    P, h, n = small_block(seed=0x1CE0A00, dt=2e-4, E=1e8)     # UpdateEveryNthStep = 25
    assert P.UpdateEveryNthStep == 25
    with_motion(P, h, n, seed=4)
    ref = h.copy()
    a = make_partition(P, h, n)
    b = make_partition(P, h, n)
    a.sort_points()
    b.sort_points()
    O.sort(P, ref, n)
    captures = []
    t, step = 0.0, 0
    for frame in range(8):
        resort = im.STEP_RESORT if frame else 0
        if frame:
            O.sort(P, ref, n)
        a.step(25, im.STEP_GRAPH | resort)
        b.step(25, resort)
        t, step, _, err = O.substeps(P, ref, n, 25, t, step)
        assert err == 0
        captures.append(a.graph_stats()[0])
    assert a.error_code == 0 and b.error_code == 0
    # 25 substeps = head + 16 + 8 + tail: four graphs per frame, keyed by the
    # grid buffer the frame starts from (three rotate; 25 substeps move it by
    # one), the clock slot (flips) and the point buffer the sort leaves
    # current (flips): the states repeat every six frames, so frames 7 and 8
    # capture nothing
    assert captures[5:] == [captures[5]] * 3 and captures[5] <= 24, captures
    out_a, out_b = download(a, h), download(b, h)
    assert_close(out_a, out_b, n, rtol=1e-10, what="graph vs direct")
    assert_close(out_a, ref, n, rtol=1e-9, what="graph vs oracle")
    a.step(1000, im.STEP_GRAPH)
    a.step(777, im.STEP_GRAPH)
    assert a.graph_stats()[1] <= 64
    assert a.error_code == 0
