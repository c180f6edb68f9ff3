"""Benchmark of the MPM substep on one MI355X.

Workload (default): the ice sheet, 200 x 12.5 m on a 4096 x 1024 grid
(h = 0.05 m), a Poisson-disk cloud of 3.97M points (4 per cell on average),
E = 9e9 Pa, nu = 0.3, rho = 980, Drucker-Prager cohesion 1e5 Pa and friction
angle 30 deg, gravity, rigid walls, a 1 m cylindrical indenter moving at 2 m/s
into the sheet, dt = 1e-5 s.  A "step" is one substep (P2G, grid update, G2P
with the return map) of every point, run as frames of UpdateEveryNthStep
substeps (500) with a device re-sort between frames, as icy::Model::Step runs
them.

Prints one JSON line:
- value: FP64 M particle-steps/s over the timed substeps (hipGraph chunk
  replays, every graph captured before the timed region);
- sort_ms_per_frame: the device re-sort of a frame boundary, timed on its own
  (inside the timed region only when it spans a frame boundary);
- roofline: the substep kernel's average launch time from hipEvents around the
  back-to-back substep launches of a separate 64-substep call (one pair per
  graph chunk), against its compulsory HBM bytes;
- mixed: the same for the mixed-precision form, with its accuracy against
  FP64 after 1000 substeps;
- batch: the 64-scene workload, mixed batched, FP64 batched and the plain FP64
  stepper scene by scene, with mixed vs FP64 accuracy after 1000 substeps;
- cpu_baseline: the FP64 CPU oracle on the 25.6k-point block (1000 substeps)
  at the host's thread count and at 1 thread, and on a sample of the sheet.

  python bench.py [--gpus 1] [--steps 2000] [--warmup 20] [--precision fp64|mixed]
  python bench.py --workload floes      # 8 floe fields, about 24M points
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "M particle-steps/s + HBM GB/s fraction, mixed vs FP64 at ≤1e-4 rel pos error"

# Compulsory HBM bytes per point and per touched grid node of the substep's
# one launch (grid update + G2P + P2G, kTileChain).  FP64: per point it reads
# utility, Jp_inv, x, Fe (8+8+16+32 = 64 B) and writes x and Fe (16+32 = 48 B):
# v and C pass to P2G in registers and are stored only by the G2P that ends a
# call, Jp_inv and the utility word only for crushed or liquid points.  Per
# node it reads the momenta of the last P2G (24 B) and forms the velocity
# while staging (the grid update folded in), accumulates mass and momentum
# into the next buffer (24 B) and clears the buffer two substeps back (24 B).
# The mixed form keeps 4-byte point fields (32 + 24 B per point) and the same
# FP64 grid (24 B per node and pass).
BYTES = {
    "fp64": {"fused_point": 64 + 48, "fused_node": 24 + 24 + 24},
    "mixed": {"fused_point": 32 + 24, "fused_node": 24 + 24 + 24},
}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sheet4m_fused_tiles_hbm_traffic.md")


def measured_traffic(precision_name, n_points):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (copy-calibrated), or None
    when the summary is absent or was taken on another point count."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    kv = dict(t.split("=") for t in line.split()[1:])
                    if int(kv.get("points", -1)) == n_points and precision_name in kv:
                        return int(float(kv[precision_name]))
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """CPU model name and logical CPU count of this host (lscpu's "Model name")."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, len(os.sched_getaffinity(0))


def oracle_rate(O, P, hssoa, n, substeps, threads):
    """M particle-steps/s of the FP64 oracle over `substeps` substeps (wall clock
    around the substeps only), after one warm-up substep."""
    O.lib().orc_set_threads(threads)
    h = hssoa.copy()
    O.sort(P, h, n)
    O.substeps(P, h, n, 1, 0.0, 0)                     # warm-up (page-in)
    t0 = time.perf_counter()
    O.substeps(P, h, n, substeps, P.InitialTimeStep, 1)
    dt = time.perf_counter() - t0
    return n * substeps / dt / 1e6, dt


def cpu_baseline(scenes, P_sheet, h_sheet, n_sheet, sheet_substeps, block_substeps):
    """The FP64 CPU oracle (test infrastructure, the checker) on the host cores.

    Main figure: the 40 x 10 m block (25,600 points, grid 256 x 128, no indenter)
    for `block_substeps` substeps at the host's thread count and at 1 thread.
    Second figure: a bounded sample of the sheet itself."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    model, ncpu = cpu_model()
    threads = int(os.environ.get("OMP_NUM_THREADS", ncpu))
    Pb, hb, nb = scenes.block_scene()
    v_all, dt_all = oracle_rate(O, Pb, hb, nb, block_substeps, threads)
    v_one, dt_one = oracle_rate(O, Pb, hb, nb, block_substeps, 1)
    out = {"value": v_all, "unit": "M particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{block_substeps} substeps of the 40 x 10 m block ({nb} points, grid 256x128, h=0.25 m, "
                     f"dt=1e-5 s), FP64 oracle (C++/OpenMP), {threads} threads, {dt_all:.1f} s wall",
           "cpu_model": model, "host_cpus": ncpu,
           "one_thread": {"value": v_one, "cores": 1, "wall_s": round(dt_one, 2)}}
    if sheet_substeps > 0:
        v_sh, dt_sh = oracle_rate(O, P_sheet, h_sheet, n_sheet, sheet_substeps, threads)
        out["sheet"] = {"value": v_sh, "cores": threads,
                        "sample": f"{sheet_substeps} substeps of the {n_sheet}-point sheet, {dt_sh:.1f} s wall"}
    O.lib().orc_set_threads(threads)
    return out


def calibration_copy():
    """A plain device-to-device copy of 1 GiB (reads 1 GiB, writes 1 GiB): the
    reference launch for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters.
    Made with the HIP runtime libicempm.so links (a second runtime in the
    process, such as torch's, does not see the GPU next to it)."""
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("calibration copy: hipMalloc failed")
    hip.hipMemset(a, 1, n)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, n, 3):          # device to device
        raise RuntimeError("calibration copy: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def field_errors(ref, mx, x_init):
    """Per-field errors of a mixed-form state against the FP64 one (both (16, n)
    arrays after the same substeps from the same start), points matched by
    position (KD-tree); x_init: the start positions (2, n) in any order."""
    from scipy.spatial import cKDTree
    n = ref.shape[1]
    _, idx = cKDTree(ref[4:6].T).query(mx[4:6].T, workers=-1)
    ok = len(np.unique(idx)) == n
    mxr = np.empty_like(mx)
    mxr[:, idx] = mx
    _, i0 = cKDTree(x_init.T).query(ref[4:6].T, workers=-1)
    x0 = x_init[:, i0]
    d = mxr[4:6] - ref[4:6]
    Fi = ref[8:12] - np.array([1, 0, 0, 1.0])[:, None]
    crushed = lambda a: (a[0].view(np.int64) & 0x10000) != 0
    return {"matched_one_to_one": bool(ok),
            "rel_pos": float(np.linalg.norm(d) / np.linalg.norm(ref[4:6])),
            "rel_disp": float(np.linalg.norm(d) / max(np.linalg.norm(ref[4:6] - x0), 1e-300)),
            "max_abs_dx_m": float(np.abs(d).max()),
            "rel_v": float(np.linalg.norm(mxr[6:8] - ref[6:8]) / max(np.linalg.norm(ref[6:8]), 1e-300)),
            "rel_F_minus_I": float(np.linalg.norm(mxr[8:12] - ref[8:12]) / max(np.linalg.norm(Fi), 1e-300)),
            "rel_Jp": float(np.linalg.norm(mxr[3] - ref[3]) / np.linalg.norm(ref[3])),
            "crushed_ref": int(crushed(ref).sum()), "crushed_mismatch": int((crushed(ref) != crushed(mxr)).sum())}


def accuracy(im, P, hssoa, n, steps):
    """Mixed vs FP64 GPU stepping of the sheet from the same state."""
    outs = {}
    for name, prec in (("fp64", im.PRECISION_FP64), ("mixed", im.PRECISION_MIXED)):
        part = im.Partition(0, 0, prec)
        part.allocate(hssoa.shape[1], P.GridXTotal)
        part.update_constants(P)
        part.transfer_points_from_soa_to_device(hssoa, 0, n)
        part.sort_points()
        part.step(steps, im.STEP_GRAPH)
        if part.error_code:
            raise RuntimeError(f"accuracy run {name}: device error code {part.error_code}")
        o = np.zeros_like(hssoa)
        part.transfer_from_device(o, 0)
        outs[name] = o[:, :n]
        part.close()
    return {"substeps": steps, **field_errors(outs["fp64"], outs["mixed"], hssoa[4:6, :n])}


def frames(step_fn, k, frame, flags, resort_flag):
    """k substeps as calls of at most one frame each, re-sorting on the device
    before every call but the first (the frame boundaries of Model::Step)."""
    done = 0
    while done < k:
        m = min(frame, k - done)
        step_fn(m, flags | (resort_flag if done > 0 else 0))
        done += m


def prime(run, captures):
    """Runs the timed schedule untimed until two passes in a row capture no
    graph: the chunk graphs depend on the grid buffer, clock slot and point
    buffer a call starts from, whose combinations recur within six calls."""
    quiet = 0
    for _ in range(12):
        before = captures()
        run()
        quiet = quiet + 1 if captures() == before else 0
        if quiet == 2:
            break


def batch_summary(args, im, scenes, sheet_fp64_value):
    """The 64-scene batch workload (log-uniform 10k-500k points per scene,
    grids 128x128 to 1024x512, per-scene E and cohesion): mixed batched, FP64
    batched and the plain FP64 stepper run scene by scene (one partition per
    scene, hipGraph replay), plus mixed vs FP64 batched accuracy."""
    t0 = time.perf_counter()
    bag = scenes.batch_scenes(n_scenes=args.scenes or 64)
    n_tot = sum(n for _, _, n in bag)
    log(f"batch: {len(bag)} scenes, {n_tot} points, setup {time.perf_counter() - t0:.1f}s")
    K = args.batch_steps
    frame = min(P.UpdateEveryNthStep for P, _, _ in bag)
    G = im.STEP_GRAPH if args.mode == "graph" else 0       # --mode eager: direct launches

    # plain: one partition per scene, stepped one after the other
    parts = []
    for P, h, n in bag:
        p = im.Partition(0, 0, im.PRECISION_FP64)
        p.allocate(h.shape[1], P.GridXTotal)
        p.update_constants(P)
        p.transfer_points_from_soa_to_device(h, 0, n)
        p.sort_points()
        parts.append(p)

    def plain_step(m, flags):
        for p in parts:
            p.step(m, flags)

    prime(lambda: frames(plain_step, K, frame, G, im.STEP_RESORT),
          lambda: sum(p.graph_stats()[0] for p in parts))
    for p in parts:
        p.synchronize()
    t0 = time.perf_counter()
    frames(plain_step, K, frame, G, im.STEP_RESORT)
    for p in parts:
        p.synchronize()
    plain = n_tot * K / (time.perf_counter() - t0) / 1e6
    if any(p.error_code for p in parts):
        raise RuntimeError("batch plain: device error code")
    for p in parts:
        p.close()

    res, states = {}, {}
    for name, prec in (("fp64", im.PRECISION_FP64), ("mixed", im.PRECISION_MIXED)):
        bt = im.Batch(0, prec)
        for P, h, n in bag:
            bt.add_scene(P, h, n)
        bt.commit()
        chosen, _ = bt.autotune(2)
        tile = bt.tile_config()
        log(f"[batch {name}] autotune: candidate {chosen} {tile}")
        prime(lambda: frames(bt.step, K, frame, G, im.STEP_RESORT), lambda: bt.graph_stats()[0])
        t0 = time.perf_counter()
        frames(bt.step, K, frame, G, im.STEP_RESORT)
        wall = time.perf_counter() - t0
        if bt.error_code:
            raise RuntimeError(f"batch {name}: device error code {bt.error_code}")
        res[name] = {"value": n_tot * K / wall / 1e6, "ms_per_step": wall / K * 1e3, "tile": tile}
        bt.close()
        # accuracy: a fresh batch from the initial state, accuracy_steps substeps, no re-sort
        if args.accuracy_steps > 0:
            bt = im.Batch(0, prec)
            for P, h, n in bag:
                bt.add_scene(P, h, n)
            bt.commit()
            bt.step(args.accuracy_steps, G | im.STEP_RESORT)
            if bt.error_code:
                raise RuntimeError(f"batch accuracy {name}: device error code {bt.error_code}")
            outs = []
            for i, (P, h, n) in enumerate(bag):
                o = np.zeros_like(h)
                bt.transfer_from_device(i, o)
                outs.append(o[:, :n])
            states[name] = outs
            bt.close()
    acc = None
    if states:
        per = [field_errors(r, m, h[4:6, :n]) for r, m, (_, h, n) in zip(states["fp64"], states["mixed"], bag)]
        ref_all = np.concatenate(states["fp64"], axis=1)
        # whole-batch relative position error (points matched per scene)
        num = sum((e["rel_pos"] * np.linalg.norm(r[4:6])) ** 2 for e, r in zip(per, states["fp64"]))
        acc = {"substeps": args.accuracy_steps,
               "rel_pos": float(np.sqrt(num) / np.linalg.norm(ref_all[4:6])),
               "max_scene_rel_pos": max(e["rel_pos"] for e in per),
               "max_scene_rel_v": max(e["rel_v"] for e in per),
               "max_scene_rel_F_minus_I": max(e["rel_F_minus_I"] for e in per),
               "crushed_mismatch": sum(e["crushed_mismatch"] for e in per),
               "matched_one_to_one": all(e["matched_one_to_one"] for e in per)}
    return {
        "workload": (f"{len(bag)} independent ice blocks, {n_tot} points, log-uniform 10k-500k per scene, "
                     "grids 128x128..1024x512, per-scene E and cohesion, dt=1e-5 s"),
        "scenes": len(bag), "points": n_tot, "substeps": K,
        "unit": "M particle-steps/s",
        "mixed_batched": round(res["mixed"]["value"], 2),
        "fp64_batched": round(res["fp64"]["value"], 2),
        "fp64_plain_per_scene": round(plain, 2),
        "mixed_batched_vs_fp64_batched": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
        "mixed_batched_vs_fp64_plain_per_scene": round(res["mixed"]["value"] / plain, 3),
        "mixed_batched_vs_fp64_sheet": round(res["mixed"]["value"] / sheet_fp64_value, 3),
        "plain_mode": ("one partition per scene, scenes in turn, each replaying its own hipGraph chunks"
                       if G else "one partition per scene, scenes in turn, direct launches"),
        "tile": {k: v["tile"] for k, v in res.items()},
        "accuracy_mixed_vs_fp64": acc,
    }


def floes_workload(args, im, scenes):
    """The floe-field workload (8 fields of seeded convex floes, 1-20 m, power
    law with exponent -2, grids 4096x1024 to 8192x2048, about 24M points),
    mixed and FP64 batched against the plain FP64 stepper scene by scene."""
    t0 = time.perf_counter()
    bag = scenes.floe_fields(n_scenes=args.scenes or 8)
    n_tot = sum(n for _, _, n in bag)
    log(f"floes: {len(bag)} scenes, {n_tot} points, setup {time.perf_counter() - t0:.1f}s")
    K = args.steps
    frame = min(P.UpdateEveryNthStep for P, _, _ in bag)
    G = im.STEP_GRAPH if args.mode == "graph" else 0       # --mode eager: direct launches
    parts = []
    for P, h, n in bag:
        p = im.Partition(0, 0, im.PRECISION_FP64)
        p.allocate(h.shape[1], P.GridXTotal)
        p.update_constants(P)
        p.transfer_points_from_soa_to_device(h, 0, n)
        p.sort_points()
        parts.append(p)

    def plain_step(m, flags):
        for p in parts:
            p.step(m, flags)

    prime(lambda: frames(plain_step, K, frame, G, im.STEP_RESORT),
          lambda: sum(p.graph_stats()[0] for p in parts))
    t0 = time.perf_counter()
    frames(plain_step, K, frame, G, im.STEP_RESORT)
    for p in parts:
        p.synchronize()
    plain = n_tot * K / (time.perf_counter() - t0) / 1e6
    for p in parts:
        p.close()
    res = {}
    for name, prec in (("fp64", im.PRECISION_FP64), ("mixed", im.PRECISION_MIXED)):
        bt = im.Batch(0, prec)
        for P, h, n in bag:
            bt.add_scene(P, h, n)
        bt.commit()
        bt.autotune(2)
        tile = bt.tile_config()
        prime(lambda: frames(bt.step, K, frame, G, im.STEP_RESORT), lambda: bt.graph_stats()[0])
        t0 = time.perf_counter()
        frames(bt.step, K, frame, G, im.STEP_RESORT)
        wall = time.perf_counter() - t0
        if bt.error_code:
            raise RuntimeError(f"floes {name}: device error code {bt.error_code}")
        res[name] = {"value": n_tot * K / wall / 1e6, "ms_per_step": wall / K * 1e3, "tile": tile}
        bt.close()
    out = {
        "metric": METRIC, "value": round(res["mixed"]["value"], 2), "unit": "M particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": 0, "ms_per_step": round(res["mixed"]["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"{len(bag)} floe fields, {n_tot} points, convex floes 1-20 m (power law -2), "
                                "grids 4096x1024..8192x2048, h=0.05 m, dt=1e-5 s, mixed form, batched, "
                                + ("hipGraph replay" if G else "direct launches")), "scenes": len(bag), "points": n_tot},
        "fp64_batched": round(res["fp64"]["value"], 2),
        "fp64_plain_per_scene": round(plain, 2),
        "mixed_batched_vs_fp64_batched": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
        "mixed_batched_vs_fp64_plain_per_scene": round(res["mixed"]["value"] / plain, 3),
        "detail": res,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the sheet length (tests only)")
    ap.add_argument("--cpu-substeps", type=int, default=1000,
                    help="substeps of the CPU baseline on the 25.6k-point block (at all threads and at 1 thread)")
    ap.add_argument("--cpu-sheet-substeps", type=int, default=100,
                    help="substeps of the CPU oracle on the sheet itself (second figure; 0 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["graph", "eager"], default="graph",
                    help="graph: frames replayed as hipGraph chunks; eager: direct launches")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision line")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch object")
    ap.add_argument("--batch-steps", type=int, default=500, help="timed substeps of the batch object")
    ap.add_argument("--accuracy-steps", type=int, default=1000, help="substeps of the mixed vs FP64 comparisons")
    ap.add_argument("--timing-substeps", type=int, default=64,
                    help="substeps of the pass with hipEvents around the launches (kernel times, roofline)")
    ap.add_argument("--calib-copy", action="store_true", help="run a 1 GiB device copy first (counter calibration)")
    ap.add_argument("--tile", default=None, help="TXxTY[xMARGIN] tile shape instead of the autotuner's choice")
    ap.add_argument("--workload", choices=["sheet", "floes"], default="sheet")
    ap.add_argument("--sampler", choices=["poisson", "jittered"], default="poisson",
                    help="point cloud of the sheet: Poisson-disk (the reference generator's) or jittered lattice")
    ap.add_argument("--scenes", type=int, default=None, help="scenes of the batch (64) or floes (8) workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU stepping (x-slab halo exchange) "
                                   "is not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    import icempm_multi_amd as im
    from icempm_multi_amd import scenes

    if args.calib_copy:
        calibration_copy()
    if args.workload == "floes":
        return floes_workload(args, im, scenes)

    nx = int(round(4000 * args.scale))
    t0 = time.perf_counter()
    P, hssoa, n = scenes.sheet_scene(nx=nx, sampler=args.sampler)
    log(f"scene: {n} points, grid {P.GridXTotal}x{P.GridY}, h={P.cellsize}, setup {time.perf_counter() - t0:.1f}s")
    nn = (P.GridXTotal + 2 * P.GridHaloSize) * P.GridY

    def timed(precision_name):
        prec = im.PRECISION_MIXED if precision_name == "mixed" else im.PRECISION_FP64
        part = im.Partition(0, 0, prec)
        part.allocate(hssoa.shape[1], P.GridXTotal)
        part.update_constants(P)
        part.transfer_points_from_soa_to_device(hssoa, 0, n)
        part.sort_points()
        if args.tile:
            t = [int(v) for v in args.tile.lower().split("x")]
            part.set_tile_shape(t[0], t[1], t[2] if len(t) > 2 else 1)
            chosen = -1
        else:
            chosen, tune_ms = part.autotune(3)
            log(f"[{precision_name}] autotune: candidate {chosen}, ms per fused launch "
                f"{['%.3f' % m for m in tune_ms]}")
        tile = part.tile_config()
        frame = P.UpdateEveryNthStep
        flags = im.STEP_GRAPH if args.mode == "graph" else 0

        frames(part.step, args.warmup, frame, flags, im.STEP_RESORT)
        prime(lambda: frames(part.step, args.steps, frame, flags, im.STEP_RESORT), lambda: part.graph_stats()[0])
        part.synchronize()
        captures0 = part.graph_stats()[0]
        t0 = time.perf_counter()
        frames(part.step, args.steps, frame, flags, im.STEP_RESORT)
        part.synchronize()
        wall = time.perf_counter() - t0
        captured_in_timed = part.graph_stats()[0] - captures0
        if part.error_code:
            raise RuntimeError(f"device error code {part.error_code}")
        # the device re-sort of a frame boundary, timed alone (host wall clock
        # around the call, which ends with the tile table built)
        sort_s = []
        for _ in range(3):
            part.synchronize()
            ts = time.perf_counter()
            part.sort_points()
            sort_s.append(time.perf_counter() - ts)
        sort_ms = 1e3 * float(np.median(sort_s))
        # kernel times: one more call with hipEvents, one pair around each run
        # of back-to-back substep launches (and around the call's first P2G and
        # last substep)
        part.reset_indenter_force_accumulator()          # also resets the kernel statistics
        part.step(args.timing_substeps, flags | im.STEP_TIMINGS | im.STEP_SPAN)
        stats = part.kernel_stats()
        log(f"[{precision_name}] kernel stats (ms, launches):", {k: (round(v[0], 3), v[1]) for k, v in stats.items()})
        part.reset_grid()
        part.p2g()
        g_act = int(np.count_nonzero(part.get_grid(0, nn)[0]))
        B = BYTES[precision_name]
        fused_ms, fused_n = stats["fused_g2p_p2g"]
        roofline = None
        if fused_n > 0:
            t_launch = fused_ms / fused_n * 1e-3
            algo = n * B["fused_point"] + g_act * B["fused_node"]
            achieved = algo / t_launch / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                        "frac": round(achieved * 1e9 / PEAK_HBM, 4),
                        "traffic": measured_traffic(precision_name, n),
                        "kernel": f"{precision_name}::k_tiles<7> (grid update + G2P + P2G)",
                        "avg_launch_us": round(t_launch * 1e6, 2), "launches_timed": fused_n,
                        "algo_bytes_per_launch": algo,
                        "bytes_per_point": B["fused_point"], "bytes_per_node": B["fused_node"],
                        "touched_nodes": g_act}
        step_s = wall / args.steps
        part.close()
        return {"value": n * args.steps / wall / 1e6, "ms_per_step": step_s * 1e3, "roofline": roofline,
                "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in stats.items() if v[1]},
                "tile": tile, "chosen": chosen, "sort_ms_per_frame": round(sort_ms, 3), "frame_substeps": frame,
                "graphs_captured_in_timed_region": captured_in_timed}

    primary = timed(args.precision)
    frame = primary["frame_substeps"]
    out = {
        "metric": METRIC, "value": round(primary["value"], 2), "unit": "M particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(primary["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"ice sheet {nx * 0.05:g} x 12.5 m, grid {P.GridXTotal}x{P.GridY}, h=0.05 m, "
                               f"{n} points ({'Poisson-disk' if args.sampler == 'poisson' else 'jittered lattice'}, 4/cell on average), "
                               f"cylindrical indenter D=1 m at 2 m/s, "
                               f"dt=1e-5 s",
                   "points": n, "grid": [P.GridXTotal, P.GridY], "precision": args.precision,
                   "tile": primary["tile"],
                   "mode": ("hipGraph chunk replay" if args.mode == "graph" else "direct launches")
                           + f", frames of {frame} substeps with a device re-sort between frames"},
        "sort_ms_per_frame": primary["sort_ms_per_frame"],
        "sort_amortised_ms_per_step": round(primary["sort_ms_per_frame"] / frame, 5),
        "graphs_captured_in_timed_region": primary["graphs_captured_in_timed_region"],
        "roofline": primary["roofline"],
        "kernel_ms": primary["kernel_ms"],
        "timing_pass": f"kernel_ms and roofline.avg_launch_us: a {args.timing_substeps}-substep call after the timed "
                       "region with hipEvents, one pair around each run of back-to-back substep launches "
                       + ("(each replayed graph chunk's, recorded inside it)" if args.mode == "graph"
                          else "(the call's, direct launches)") + " and one around its first P2G and last substep",
    }
    if args.precision == "fp64" and not args.no_mixed:
        mx = timed("mixed")
        acc = accuracy(im, P, hssoa, n, args.accuracy_steps) if args.accuracy_steps > 0 else None
        out["mixed"] = {"value": round(mx["value"], 2), "unit": "M particle-steps/s",
                        "speedup_vs_fp64": round(mx["value"] / primary["value"], 3),
                        "ms_per_step": round(mx["ms_per_step"], 4), "roofline": mx["roofline"],
                        "kernel_ms": mx["kernel_ms"], "tile": mx["tile"],
                        "sort_ms_per_frame": mx["sort_ms_per_frame"], "accuracy_vs_fp64": acc}
    if args.precision == "fp64" and not args.no_batch and args.scale == 1.0:
        out["batch"] = batch_summary(args, im, scenes, primary["value"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scenes, P, hssoa, n, args.cpu_sheet_substeps, args.cpu_substeps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
