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

# algorithmic bytes: per point and per touched node, per launch
BYTES = {
    "fp64": {"fused_point": 64 + 112, "fused_node": 16 + 24, "p2g_point": 104, "g2p_point": 64 + 112,
             "step_point": 280, "step_node": 80},
    "mixed": {"fused_point": 32 + 56, "fused_node": 8 + 24, "p2g_point": 52, "g2p_point": 32 + 56,
              "step_point": 140, "step_node": 64},
}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the sheet length (tests only)")
    ap.add_argument("--cpu-substeps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay frames as hipGraphs (no per-kernel events)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU stepping (x-slab halo exchange) "
                                   "is not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    import icempm_multi_amd as im
    from icempm_multi_amd import scenes

    nx = int(round(4000 * args.scale))
    t0 = time.perf_counter()
    P, hssoa, n = scenes.sheet_scene(nx=nx)
    log(f"scene: {n} points, grid {P.GridXTotal}x{P.GridY}, h={P.cellsize}, setup {time.perf_counter() - t0:.1f}s")
    prec = im.PRECISION_MIXED if args.precision == "mixed" else im.PRECISION_FP64
    part = im.Partition(0, 0, prec)
    part.allocate(hssoa.shape[1], P.GridXTotal)
    part.update_constants(P)
    part.transfer_points_from_soa_to_device(hssoa, 0, n)
    part.sort_points()
    chosen, tune_ms = part.autotune(3)
    log(f"autotune: candidate {chosen}, ms/substep {['%.3f' % m for m in tune_ms[:8]]}")

    frame = P.UpdateEveryNthStep
    flags_timed = im.STEP_TIMINGS if not args.graph else im.STEP_GRAPH

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
    value = n * args.steps / wall / 1e6
    stats = part.kernel_stats()
    log("kernel stats (ms, launches):", {k: (round(v[0], 2), v[1]) for k, v in stats.items()})

    # nodes touched by the points (for the grid part of the algorithmic bytes)
    nn = (P.GridXTotal + 2 * P.GridHaloSize) * P.GridY
    part.reset_grid()
    part.p2g()
    g_act = int(np.count_nonzero(part.get_grid(0, nn)[0]))
    B = BYTES[args.precision]
    fused_ms, fused_n = stats["fused_g2p_p2g"]
    roofline = None
    if fused_n > 0:
        t_launch = fused_ms / fused_n * 1e-3
        algo = n * B["fused_point"] + g_act * B["fused_node"]
        achieved = algo / t_launch / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": round(achieved * 1e9 / PEAK_HBM, 4), "traffic": None,
                    "kernel": "k_tiles<fused G2P+P2G>", "avg_launch_us": round(t_launch * 1e6, 2),
                    "algo_bytes_per_launch": algo}
    step_bytes = n * B["step_point"] + P.GridXTotal * P.GridY * B["step_node"]
    step_s = wall / args.steps
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "M particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"ice sheet {nx * 0.05:g} x 12.5 m, grid {P.GridXTotal}x{P.GridY}, h=0.05 m, "
                               f"{n} points (4/cell jittered), cylindrical indenter D=1 m at 2 m/s, dt=1e-5 s",
                   "points": n, "grid": [P.GridXTotal, P.GridY], "precision": args.precision,
                   "tile": chosen, "mode": "graph" if args.graph else "eager+events"},
        "hbm_model_GBps": round(step_bytes / step_s / 1e9, 1),
        "hbm_model_frac": round(step_bytes / step_s / PEAK_HBM, 4),
        "roofline": roofline,
        "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in stats.items() if v[1]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, hssoa, n, args.cpu_substeps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
