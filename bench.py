"""Benchmark: plain GPU DMO on the 40^3 x 6 tet cube (vertex updates/s).

A step is one DMO iteration over the whole mesh (every inner colour class
searched once: 59,319 vertex searches of 8^3 candidates x 3 levels).  The
mesh is built and resident in HBM before the timed region; W warm-up
iterations run on a separate copy of the mesh, then K iterations of a fresh
mesh are timed (wall clock around the blocking iterate call, which returns
after the last kernel and the per-iteration quality check).

Output: one JSON line with the metric, a roofline object for the dominant
kernel (tet search; algorithmic FLOP per vertex update x updates per launch /
mean launch duration from HIP events on the launch stream) and a CPU baseline
(the oracle, a C/OpenMP restatement, on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex updates/s at equal final min mean-ratio quality; learned vs plain DMO"
# algorithmic work of one tet vertex update (8^3 candidates + current position
# per level, 3 levels, plus the final check: 1,540 ring evaluations x 24 tets
# x 73 FLOP, cube roots counted separately)
FLOP_PER_UPDATE_TET = 2_698_080
CBRT_PER_UPDATE_TET = 36_960
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per launch of the tet search on this workload: rocprofv3 --pmc
# FETCH_SIZE 5,007.5 KiB x 2.0 (copy-kernel calibration) + WRITE_SIZE 694.9 KiB
# x 1.0, mean over the 24 dispatches of 3 iterations
# (profiles/k_search_tet_tet40.md)
TRAFFIC_BYTES_PER_LAUNCH = 10_966_938


def cpu_baseline(xyz, tets, iters: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    m = oracle_py.OracleMesh("tet", xyz, tets)
    n_inner = m.m.nInner
    t0 = time.perf_counter()
    for _ in range(iters):
        m.iterate(nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": iters * n_inner / dt, "unit": "vertex updates/s", "cores": threads, "kind": "port",
            "sample": f"{iters} full DMO iteration(s) of the same 40^3 x 6 tet cube ({iters * n_inner} vertex "
                      f"searches, {dt:.1f} s) by the C/OpenMP oracle with {threads} threads"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2)
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this DMO build runs on a single MI355X and has no multi-GPU "
              "path; run it with --gpus 1.", file=sys.stderr)
        return 2

    from volumesurfacediscretemeshoptimization_amd import DMOMeshTetFactory, DMOTetClass, synth

    xyz, tets = synth.tet_cube(args.cells, seed=args.seed)
    # warm-up on a throwaway copy
    if args.warmup > 0:
        wm = DMOMeshTetFactory.create(xyz, tets)
        wo = DMOTetClass(wm, n_iter=args.warmup)
        wo.iterate(args.warmup)
        wo.close()
        wm.close()
    t0 = time.perf_counter()
    mesh = DMOMeshTetFactory.create(xyz, tets)
    build_s = time.perf_counter() - t0
    counts = mesh.counts()
    opt = DMOTetClass(mesh, n_iter=args.steps)
    q0 = opt.getQuality()
    opt.set_timing(True)
    t0 = time.perf_counter()
    opt.iterate(args.steps)
    dt = time.perf_counter() - t0
    st = opt.stats()
    q1 = opt.getQuality()
    qmean = opt.getMeanQuality()
    updates = st["searches_inner"]
    value = updates / dt
    flops = updates * FLOP_PER_UPDATE_TET
    kernel_s = st["kernel_ms"] / 1e3
    launches = st["kernel_launches"]
    achieved = flops / kernel_s / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "kernel": "k_search_tet<8>", "launches": launches,
            "mean_launch_us": 1e3 * st["kernel_ms"] / max(launches, 1),
            "flop_per_update": FLOP_PER_UPDATE_TET, "cbrt_per_update": CBRT_PER_UPDATE_TET}
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(xyz, tets, args.cpu_iters, threads)
    out = {
        "metric": METRIC, "value": value, "unit": "vertex updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain GPU DMO, tet cube {args.cells}^3 cells x 6 Kuhn tets "
                               f"({counts['n_vertices']} vertices, {counts['n_elements']} tets), interior jitter "
                               f"+-0.35h seed {args.seed}, boundary fixed; 8^3 candidates x 3 levels per vertex; "
                               "one step = one iteration over all inner colour classes",
                   "vertex_updates_per_step": counts["n_inner"], "colors_inner": counts["n_colors_inner"],
                   "mesh_build_s": build_s, "min_quality_start": q0, "min_quality_end": q1,
                   "mean_quality_end": qmean},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    opt.close()
    mesh.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
