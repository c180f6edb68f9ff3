"""Benchmark of the MPM substep on one MI355X.

Workload (default): the ice sheet, 200 x 12.5 m on a 4096 x 1024 grid
(h = 0.05 m), 4 jittered points per cell (4.0M points), E = 9e9 Pa, nu = 0.3,
rho = 980, Drucker-Prager cohesion 1e5 Pa and friction angle 30 deg, gravity,
rigid walls, a 1 m cylindrical indenter moving at 2 m/s into the sheet, dt =
1e-5 s.  A "step" is one substep (reset, P2G, grid update, G2P with the
return map) of every point; points are re-binned on the device at every frame
boundary (UpdateEveryNthStep substeps), inside the timed region.

Prints one JSON line: whole-job throughput in M particle-steps/s, the HBM
roofline of the dominant kernel (the fused G2P+P2G launch) from HIP events over
the timed region, and the FP64 CPU oracle timed on a bounded sample of the same
workload on the host cores.

  python bench.py [--gpus 1] [--steps 2000] [--warmup 20] [--precision fp64|mixed]
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

# Algorithmic HBM bytes per point and per touched grid node, per launch.
# FP64 fused G2P+P2G: reads utility, Jp_inv, x, Fe (8+8+16+32 = 64 B) and writes
# x and Fe (16+32 = 48 B) per point: v and C pass to P2G in registers and are
# stored only by the plain G2P launch that ends a call, Jp_inv and the utility
# word only for crushed or liquid points.  Per node it reads the velocity (16 B)
# and accumulates mass and momentum (24 B).  The mixed form keeps 4-byte point
# fields (32 + 24 B); its node velocity is 8 B and the compensated (hi, lo)
# mass/momentum pairs 24 B.  A whole substep adds the grid update (per node:
# read 3 values, write 2, clear the other buffer).
BYTES = {
    "fp64": {"fused_point": 64 + 48, "fused_node": 16 + 24, "p2g_point": 104, "g2p_point": 64 + 112,
             "step_point": 280, "step_node": 80},
    "mixed": {"fused_point": 32 + 24, "fused_node": 8 + 24, "p2g_point": 52, "g2p_point": 32 + 56,
              "step_point": 140, "step_node": 64},
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


def cpu_baseline(P, hssoa, n, substeps):
    """The FP64 CPU oracle (test infrastructure, the checker) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.lib().orc_set_threads(threads)
    h = hssoa.copy()
    O.substeps(P, h, n, 1, 0.0, 0)                     # warm-up (page-in)
    t0 = time.perf_counter()
    O.substeps(P, h, n, substeps, 1e-5, 1)
    dt = time.perf_counter() - t0
    return {"value": n * substeps / dt / 1e6, "unit": "M particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{substeps} substeps of the same {n}-point sheet, FP64 oracle (C++/OpenMP), "
                      f"{dt:.1f} s wall"}


def calibration_copy():
    """A plain device-to-device copy of 1 GiB (reads 1 GiB, writes 1 GiB): the
    reference launch for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def accuracy(im, P, hssoa, n, steps):
    """Mixed vs FP64 GPU stepping from the same state, points matched by
    position (KD-tree): relative position error and per-field errors."""
    from scipy.spatial import cKDTree
    outs = {}
    for name, prec in (("fp64", im.PRECISION_FP64), ("mixed", im.PRECISION_MIXED)):
        part = im.Partition(0, 0, prec)
        part.allocate(hssoa.shape[1], P.GridXTotal)
        part.update_constants(P)
        part.transfer_points_from_soa_to_device(hssoa, 0, n)
        part.sort_points()
        part.step(steps, im.STEP_GRAPH)
        o = np.zeros_like(hssoa)
        part.transfer_from_device(o, 0)
        outs[name] = o[:, :n]
        part.close()
    ref, mx = outs["fp64"], outs["mixed"]
    _, idx = cKDTree(ref[4:6].T).query(mx[4:6].T, workers=-1)
    ok = len(np.unique(idx)) == n
    mxr = np.empty_like(mx)
    mxr[:, idx] = mx
    x0 = np.empty((2, n))
    _, i0 = cKDTree(hssoa[4:6, :n].T).query(ref[4:6].T, workers=-1)
    x0[:] = hssoa[4:6, :n][:, i0]
    d = mxr[4:6] - ref[4:6]
    Fi = ref[8:12] - np.array([1, 0, 0, 1.0])[:, None]
    crushed = lambda a: (a[0].view(np.int64) & 0x10000) != 0
    return {"substeps": steps, "matched_one_to_one": bool(ok),
            "rel_pos": float(np.linalg.norm(d) / np.linalg.norm(ref[4:6])),
            "rel_disp": float(np.linalg.norm(d) / max(np.linalg.norm(ref[4:6] - x0), 1e-300)),
            "max_abs_dx_m": float(np.abs(d).max()),
            "rel_v": float(np.linalg.norm(mxr[6:8] - ref[6:8]) / max(np.linalg.norm(ref[6:8]), 1e-300)),
            "rel_F_minus_I": float(np.linalg.norm(mxr[8:12] - ref[8:12]) / max(np.linalg.norm(Fi), 1e-300)),
            "rel_Jp": float(np.linalg.norm(mxr[3] - ref[3]) / np.linalg.norm(ref[3])),
            "crushed_ref": int(crushed(ref).sum()), "crushed_mismatch": int((crushed(ref) != crushed(mxr)).sum())}


def batch_workload(args, im, scenes):
    """Many independent scenes, FP64 and mixed, stepped as one launch per phase,
    against stepping each scene with its own partition.

    batch: 64 blocks (log-uniform 10k-500k points, grids 128x128 to 1024x512,
    E ~ U[3e9, 9e9], cohesion ~ U[5e4, 2e5]).
    floes: 8 fields of seeded convex floes (diameters 1-20 m, power law with
    exponent -2) on grids of 4096x1024 to 8192x2048, about 24M points."""
    t0 = time.perf_counter()
    if args.workload == "floes":
        bag = scenes.floe_fields(n_scenes=args.scenes)
        n_tot = sum(n for _, _, n in bag)
        desc = (f"{len(bag)} floe fields, {n_tot} points, convex floes 1-20 m (power law -2), grids "
                "4096x1024..8192x2048, h=0.05 m, dt=1e-5 s, hipGraph replay")
    else:
        bag = scenes.batch_scenes(n_scenes=args.scenes)
        n_tot = sum(n for _, _, n in bag)
        desc = (f"{len(bag)} independent ice blocks, {n_tot} points, log-uniform 10k-500k per scene, "
                "grids 128x128..1024x512, per-scene E and cohesion, dt=1e-5 s, hipGraph replay")
    log(f"{args.workload}: {len(bag)} scenes, {n_tot} points, setup {time.perf_counter() - t0:.1f}s")
    K = args.steps

    # plain: one partition per scene, scenes stepped one after the other
    parts = []
    for P, h, n in bag:
        p = im.Partition(0, 0, im.PRECISION_FP64)
        p.allocate(h.shape[1], P.GridXTotal)
        p.update_constants(P)
        p.transfer_points_from_soa_to_device(h, 0, n)
        p.sort_points()
        p.step(2, 0)
        parts.append(p)
    for p in parts:
        p.synchronize()
    t0 = time.perf_counter()
    for p in parts:
        p.step(K, 0)
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
        chosen, ms = bt.autotune(2)
        log(f"[{args.workload} {name}] autotune: candidate {chosen}, ms/substep {['%.3f' % m for m in ms[:28]]}")
        bt.reset_kernel_stats()
        bt.step(20, im.STEP_TIMINGS)
        st = bt.kernel_stats()
        log(f"[{args.workload} {name}] kernel ms per launch: "
            f"{ {k: round(v[0] / v[1], 4) for k, v in st.items() if v[1]} }")
        bt.step(args.warmup, im.STEP_GRAPH)
        bt.step(K, im.STEP_GRAPH)          # captures the K-substep graph
        t0 = time.perf_counter()
        bt.step(K, im.STEP_GRAPH)          # replay
        wall = time.perf_counter() - t0
        if bt.error_code:
            raise RuntimeError(f"batch {name}: device error code {bt.error_code}")
        res[name] = {"value": n_tot * K / wall / 1e6, "ms_per_step": wall / K * 1e3, "tile": chosen,
                     "device_ms_per_step": bt.last_step_ms()}
        bt.close()
    out = {
        "metric": METRIC, "value": round(res["mixed"]["value"], 2), "unit": "M particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(res["mixed"]["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": desc, "scenes": len(bag), "points": n_tot},
        "fp64_batched": round(res["fp64"]["value"], 2),
        "fp64_plain_per_scene": round(plain, 2),
        "speedup_mixed_batched_vs_plain_fp64": round(res["mixed"]["value"] / plain, 3),
        "speedup_fp64_batched_vs_plain_fp64": round(res["fp64"]["value"] / plain, 3),
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
    ap.add_argument("--cpu-substeps", type=int, default=300, help="substeps of the CPU baseline sample (~15 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["graph", "eager"], default="graph",
                    help="graph: frames replayed as hipGraphs with hipEvents recorded inside the graph around "
                         "the launches of one substep in 32; eager: direct launches with events around each")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision line")
    ap.add_argument("--accuracy-steps", type=int, default=1000, help="substeps of the mixed vs FP64 comparison")
    ap.add_argument("--calib-copy", action="store_true", help="run a 1 GiB device copy first (counter calibration)")
    ap.add_argument("--tile", default=None, help="TXxTY tile shape instead of the autotuner's choice")
    ap.add_argument("--workload", choices=["sheet", "batch", "floes"], default="sheet")
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
    if args.workload in ("batch", "floes"):
        if args.scenes is None:
            args.scenes = 64 if args.workload == "batch" else 8
        return batch_workload(args, im, scenes)

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
            tx, ty = (int(v) for v in args.tile.lower().split("x"))
            part.set_tile_shape(tx, ty)
            chosen = args.tile
        else:
            chosen, tune_ms = part.autotune(3)
            log(f"[{precision_name}] autotune: candidate {chosen}, ms/substep {['%.3f' % m for m in tune_ms[:28]]}")
        frame = P.UpdateEveryNthStep
        flags_timed = (im.STEP_GRAPH | im.STEP_TIMINGS | im.STEP_SAMPLED) if args.mode == "graph" else im.STEP_TIMINGS

        def run(k, flags):
            done = 0
            while done < k:
                m = min(frame - (part.clock[1] % frame), k - done)
                resort = im.STEP_RESORT if part.clock[1] % frame == 0 and part.clock[1] > 0 else 0
                part.step(m, flags | resort)
                done += m

        run(args.warmup, flags_timed)
        part.synchronize()
        part.reset_indenter_force_accumulator()          # also resets the kernel statistics
        t0 = time.perf_counter()
        run(args.steps, flags_timed)
        part.synchronize()
        wall = time.perf_counter() - t0
        if part.error_code:
            raise RuntimeError(f"device error code {part.error_code}")
        stats = part.kernel_stats()
        log(f"[{precision_name}] kernel stats (ms, launches):", {k: (round(v[0], 2), v[1]) for k, v in stats.items()})
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
                        "kernel": f"{precision_name}::k_tiles<3> (fused G2P+P2G)",
                        "avg_launch_us": round(t_launch * 1e6, 2), "algo_bytes_per_launch": algo,
                        "bytes_per_point": B["fused_point"], "bytes_per_node": B["fused_node"],
                        "touched_nodes": g_act}
        step_s = wall / args.steps
        step_bytes = n * B["step_point"] + P.GridXTotal * P.GridY * B["step_node"]
        part.close()
        return {"value": n * args.steps / wall / 1e6, "ms_per_step": step_s * 1e3, "roofline": roofline,
                "hbm_model_GBps": step_bytes / step_s / 1e9, "hbm_model_frac": step_bytes / step_s / PEAK_HBM,
                "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in stats.items() if v[1]}, "tile": chosen}

    primary = timed(args.precision)
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
                   "tile": primary["tile"], "mode": "graph, events in 1 of 32 substeps" if args.mode == "graph" else "eager, events per launch"},
        "hbm_model_GBps": round(primary["hbm_model_GBps"], 1),
        "hbm_model_frac": round(primary["hbm_model_frac"], 4),
        "roofline": primary["roofline"],
        "kernel_ms": primary["kernel_ms"],
    }
    if args.precision == "fp64" and not args.no_mixed:
        mx = timed("mixed")
        acc = accuracy(im, P, hssoa, n, args.accuracy_steps) if args.accuracy_steps > 0 else None
        out["mixed"] = {"value": round(mx["value"], 2), "unit": "M particle-steps/s",
                        "speedup_vs_fp64": round(mx["value"] / primary["value"], 3),
                        "ms_per_step": round(mx["ms_per_step"], 4), "roofline": mx["roofline"],
                        "hbm_model_frac": round(mx["hbm_model_frac"], 4), "kernel_ms": mx["kernel_ms"],
                        "accuracy_vs_fp64": acc}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, hssoa, n, args.cpu_substeps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
