"""The C++ host layer on the GPU: icy::Model::Step frames driven through
GPU_Implementation5 / GPU_Partition and the C-ABI (icempm_run), on the
reference's p1 configuration and 2d_70k input, against the FP64 oracle.

Bars: FP64 positions/velocities/F within rtol 1e-8 of the oracle after two
frames (1000 substeps) in both the device-loop and the host-loop modes; the
frame-averaged indenter force of indenter_force.csv within 1e-8 of the
oracle's; mixed precision rel. position error <= 1e-4; a snapshot written
after frame 1 and resumed for frame 2 ends where two straight frames end.
"""
import os

import numpy as np
import pytest

import _host as H
import _oracle as O
import icempm_multi_amd as im
from icempm_multi_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = {"x": (4, 6), "v": (6, 8), "F": (8, 12), "C": (12, 16), "Jp": (3, 4)}


def inputs(tmp):
    d = np.load(os.path.join(H.GOLDEN, "ref_2d_70k.npz"))
    H.write_flat_raw(tmp / "in.bin", d["x"], d["y"], d["llGrainIDs"], d["volume"][0])
    H.run("--convert-raw", tmp / "in.bin", tmp / "pts.h5")
    cfg = H.p1_config(str(tmp))
    P, _ = im.params_parse_json(open(cfg).read())
    h, n = scenes.load_raw_points(P, d["x"], d["y"], d["llGrainIDs"], float(d["volume"][0]))
    return cfg, P, h, n


def oracle_frames(P, h, n, frames, resort=True):
    """Frames of UpdateEveryNthStep substeps; a stable re-sort before each
    frame as the device loop does.  Returns state and per-frame mean force."""
    ref = h.copy()
    t, step = 0.0, 0
    forces = []
    for _ in range(frames):
        if resort:
            O.sort(P, ref, n)
        t, step, f, err = O.substeps(P, ref, n, P.UpdateEveryNthStep, t, step)
        assert err == 0
        forces.append(f.reshape(-1, 2).sum(0) / P.UpdateEveryNthStep)
    return ref, forces


def assert_close(out, ref, n, rtol, what):
    for name, (a, b) in FIELDS.items():
        s = max(np.abs(ref[a:b, :n]).max(), 1e-300)
        d = np.abs(out[a:b, :n] - ref[a:b, :n]).max()
        assert d <= rtol * s, f"{what} {name}: max diff {d:.3e} scale {s:.3e}"
    assert np.array_equal(out[0, :n].view(np.int64), ref[0, :n].view(np.int64)), f"{what}: utility words"


def read_force_log(path):
    rows = [list(map(float, l.split(","))) for l in open(path).read().strip().splitlines()]
    return np.array(rows)   # t, |f|, fx, fy


@pytest.mark.parametrize("mode", ["device", "host"])
def test_model_frames_fp64_match_oracle(tmp_path, mode):
    cfg, P, h, n = inputs(tmp_path)
    extra = ["--host-loop"] if mode == "host" else []
    out_s = H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 2, "--log-dir", tmp_path / "logs",
                  "--dump-soa", tmp_path / "out.bin", *extra, cwd=tmp_path)
    assert "done: step 1000" in out_s
    out, size = H.read_soa_dump(tmp_path / "out.bin")
    assert size == n
    # the host loop does not re-sort; the device loop re-sorts each frame
    ref, forces = oracle_frames(P, h, n, 2, resort=(mode == "device"))
    assert_close(out, ref, n, 1e-8, f"model/{mode}")
    log = read_force_log(tmp_path / "logs" / "indenter_force.csv")
    assert log.shape == (2, 4)
    assert np.allclose(log[:, 0], [500 * 1e-5, 1000 * 1e-5], rtol=1e-12)
    for k in range(2):
        assert np.allclose(log[k, 2:], forces[k], rtol=1e-8, atol=1e-8 * np.abs(forces[k]).max())
    timings = open(tmp_path / "logs" / "timings.csv").read().strip().splitlines()
    assert len(timings) == 2 and float(timings[-1].split(",")[-1]) > 0


def test_model_rebalance_path_matches_oracle(tmp_path):
    """ExtraSpaceForIncomingPoints = 0.1 leaves less free point space than
    RebalanceThresholdFreeSpaceRemaining (0.15), so every frame ends with the
    rebalance of model.cpp:102-110.  With one partition it runs on the device
    (compaction and sort, the host copy permuted to match): the points are
    uploaded once, at load, and the run follows the oracle."""
    d = np.load(os.path.join(H.GOLDEN, "ref_2d_70k.npz"))
    H.write_flat_raw(tmp_path / "in.bin", d["x"], d["y"], d["llGrainIDs"], d["volume"][0])
    H.run("--convert-raw", tmp_path / "in.bin", tmp_path / "pts.h5")
    cfg = H.p1_config(str(tmp_path), ExtraSpaceForIncomingPoints=0.1)
    P, _ = im.params_parse_json(open(cfg).read())
    h, n = scenes.load_raw_points(P, d["x"], d["y"], d["llGrainIDs"], float(d["volume"][0]))
    out_s = H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 2, "--log-dir", "",
                  "--dump-soa", tmp_path / "out.bin", cwd=tmp_path)
    assert out_s.count("rebalancing done") == 2
    assert out_s.count("rebalancing on device") == 2
    assert "uploads 1 device_rebalances 2" in out_s
    out, size = H.read_soa_dump(tmp_path / "out.bin")
    assert size == n
    ref, _ = oracle_frames(P, h, n, 2)
    # the device sort after frame 2 orders the points as the oracle's sort does
    O.sort(P, ref, n)
    assert_close(out, ref, n, 1e-8, "rebalance")


def test_device_rebalance_drops_disabled_points(tmp_path):
    """Points disabled in the input are dropped by the device rebalance, and
    the host copy keeps exactly the device's points in the device's order: the
    run equals the oracle on the input with the disabled points removed."""
    d = np.load(os.path.join(H.GOLDEN, "ref_2d_70k.npz"))
    H.write_flat_raw(tmp_path / "in.bin", d["x"], d["y"], d["llGrainIDs"], d["volume"][0])
    H.run("--convert-raw", tmp_path / "in.bin", tmp_path / "pts.h5")
    cfg = H.p1_config(str(tmp_path), ExtraSpaceForIncomingPoints=0.1)
    P, _ = im.params_parse_json(open(cfg).read())
    h, n = scenes.load_raw_points(P, d["x"], d["y"], d["llGrainIDs"], float(d["volume"][0]))
    # --no-resort: the frame runs on the loaded order with the disabled points
    # in place, and the rebalance at its end is the first device sort
    out_s = H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 1, "--log-dir", "", "--no-resort",
                  "--dump-soa", tmp_path / "out.bin", "--disable-every", 97, cwd=tmp_path)
    dropped = len(range(0, n, 97))
    assert f"{n - dropped} points kept, {dropped} disabled dropped" in out_s
    assert "uploads 2 device_rebalances 1" in out_s          # load, re-upload with the flags; none after
    out, size = H.read_soa_dump(tmp_path / "out.bin")
    assert size == n - dropped
    ref = h.copy()
    O.sort(P, ref, n)                                      # the order LoadRawPoints leaves
    ud = ref[0, :n].view(np.int64)
    ud[::97] |= 0x20000
    ref, _ = oracle_frames(P, ref, n, 1, resort=False)
    m, _, _ = O.sort(P, ref, n)
    assert m == n - dropped
    assert_close(out, ref, m, 1e-8, "rebalance with disabled points")


def test_model_frames_mixed_accuracy(tmp_path):
    cfg, P, h, n = inputs(tmp_path)
    H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 2, "--precision", "mixed", "--log-dir", "",
          "--dump-soa", tmp_path / "out.bin", "--quiet", cwd=tmp_path)
    out, size = H.read_soa_dump(tmp_path / "out.bin")
    ref, _ = oracle_frames(P, h, n, 2)
    X, R = out[4:6, :n], ref[4:6, :n]
    rel = np.linalg.norm(X - R) / np.linalg.norm(R)
    print("2d_70k/p1 mixed, 1000 substeps: rel_pos", rel)
    assert rel <= 1e-4


def test_snapshot_resume_continues_the_run(tmp_path):
    cfg, P, h, n = inputs(tmp_path)
    H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 2, "--log-dir", "", "--quiet",
          "--dump-soa", tmp_path / "straight.bin", cwd=tmp_path)
    H.run(cfg, "--points", tmp_path / "pts.h5", "--frames", 1, "--log-dir", "", "--quiet",
          "--snapshot-dir", tmp_path / "snap", cwd=tmp_path)
    snap = tmp_path / "snap" / "snapshots" / "d00001.h5"
    assert snap.exists()
    out_s = H.run(cfg, "--resume", snap, "--frames", 1, "--log-dir", "", "--quiet",
                  "--dump-soa", tmp_path / "resumed.bin", cwd=tmp_path)
    assert "done: step 1000" in out_s
    a, na = H.read_soa_dump(tmp_path / "straight.bin")
    b, nb = H.read_soa_dump(tmp_path / "resumed.bin")
    assert na == nb == n
    assert_close(b, a, n, 1e-9, "resumed")
