"""Benchmark: plain GPU DMO (vertex updates/s).

Default workload: the 40^3 x 6 tet cube (68,921 vertices, 384,000 tets,
interior jitter +-0.35h, boundary fixed, 8^3 candidates x 3 levels).  A step
is one DMO iteration over the whole mesh: every colour class searched once
(59,319 inner vertex searches for the tet cube).  Other workloads
(--workload): hex48 = 48^3 hex cube with the surface free (117,649
vertices), tri64 = 64 x 64 flat triangle grid (boundary fixed).

The mesh is built and resident in HBM before the timed region; W warm-up
iterations run on a separate copy of the mesh, then K iterations of a fresh
mesh are timed (wall clock around the blocking iterate call, which returns
after the last kernel and the per-iteration quality check).

Output: one JSON line with the metric, a roofline object for the dominant
kernel (the inner search; algorithmic FLOP per vertex update x updates per
launch / mean launch duration from HIP events on the launch stream) and a
CPU baseline (the oracle, a C/OpenMP restatement, on the host cores).
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
PEAK_FP32_TFLOPS = 157.3
# Algorithmic work of one vertex update: (nq^d + 1) ring evaluations per level
# x levels + the final check = 1,540 (3D) / 196 (2D), times the one-ring size
# (24 tets / 8 hexes / 6 triangles) times the per-element FLOP of the metric
# (tet 73, hex 407, tri 23), cube roots counted separately.
# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE x 2.0 copy-kernel
# calibration + WRITE_SIZE x 1.0, mean over the dispatches of 3 iterations,
# profiles/k_search_tet_tet40.md and profiles/k_search_hex_hex48.md).
WORKLOADS = {
    "tet40": {"kind": "tet", "flop": 2_698_080, "cbrt": 36_960, "kernel": "k_search_tet<8>",
              "traffic": 10_966_938},
    "hex48": {"kind": "hex", "flop": 5_014_240, "cbrt": 110_880, "kernel": "k_search_hex<8>",
              "traffic": None},
    "tri64": {"kind": "tri", "flop": 27_048, "cbrt": 0, "kernel": "k_search_flat<8>", "traffic": None},
}


def make_mesh(name: str, seed: int):
    from volumesurfacediscretemeshoptimization_amd import synth
    if name == "tet40":
        xyz, cells = synth.tet_cube(40, seed=seed)
        desc = ("plain GPU DMO, tet cube 40^3 cells x 6 Kuhn tets (68,921 vertices, 384,000 tets), interior "
                f"jitter +-0.35h seed {seed}, boundary fixed; 8^3 candidates x 3 levels per vertex; one step = "
                "one iteration over all inner colour classes")
        return xyz, cells, desc, {"fixed_boundary": True}
    if name == "hex48":
        xyz, cells = synth.hex_cube(48, seed=seed)
        desc = ("plain GPU DMO, hex cube 48^3 cells (117,649 vertices, 110,592 hexes), interior jitter +-0.3h, "
                f"surface vertices +-0.3h tangentially, seed {seed}, surface free (quadric fits; surface passes "
                "run only while the minimum quality is > 0); 8^3 candidates x 3 levels per inner vertex; one "
                "step = one iteration over all colour classes")
        return xyz, cells, desc, {"fixed_boundary": False}
    if name == "tri64":
        xyz, cells = synth.tri_grid(64, seed=seed)
        desc = ("plain GPU DMO, flat triangle grid 64 x 64 vertices (7,938 triangles), interior jitter +-0.4h, "
                f"seed {seed}, boundary fixed; 8 x 8 candidates x 3 levels; one step = one iteration")
        return xyz, cells, desc, {"fixed_boundary": True}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(kind, xyz, cells, iters: int, threads: int, surface: bool, name: str):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    m = oracle_py.OracleMesh(kind, xyz, cells)
    if surface:
        m.surface_init()
    t0 = time.perf_counter()
    n = 0
    for _ in range(iters):
        if surface:
            n2, n1 = m.iterate_surface(nthreads=threads)
            n += m.m.nInner + n2 + n1
        else:
            n += m.iterate(nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "vertex updates/s", "cores": threads, "kind": "port",
            "sample": f"{iters} full DMO iteration(s) of the same {name} mesh ({n} vertex searches, {dt:.1f} s) "
                      f"by the C/OpenMP oracle with {threads} threads"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="tet40")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2)
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this DMO build runs on a single MI355X and has no multi-GPU "
              "path; run it with --gpus 1.", file=sys.stderr)
        return 2

    from volumesurfacediscretemeshoptimization_amd import (DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                           DMOMeshTriFactory, DMOTetClass, DMOTriFlatClass)

    wl = WORKLOADS[args.workload]
    kind = wl["kind"]
    fac = {"tet": DMOMeshTetFactory, "hex": DMOMeshHexFactory, "tri": DMOMeshTriFactory}[kind]
    cls = {"tet": DMOTetClass, "hex": DMOHexClass, "tri": DMOTriFlatClass}[kind]
    xyz, cells, desc, kw = make_mesh(args.workload, args.seed)
    # warm-up on a throwaway copy
    if args.warmup > 0:
        wm = fac.create(xyz, cells)
        wo = cls(wm, n_iter=args.warmup, **kw)
        wo.iterate(args.warmup)
        wo.close()
        wm.close()
    t0 = time.perf_counter()
    mesh = fac.create(xyz, cells)
    build_s = time.perf_counter() - t0
    counts = mesh.counts()
    opt = cls(mesh, n_iter=args.steps, **kw)
    q0 = opt.getQuality()
    opt.set_timing(True)
    t0 = time.perf_counter()
    opt.iterate(args.steps)
    dt = time.perf_counter() - t0
    st = opt.stats()
    q1 = opt.getQuality()
    qmean = opt.getMeanQuality()
    searched = st["searches_inner"] if kind != "tri" else st["searches_surface"]
    updates = st["searches_inner"] + st["searches_surface"] + st["searches_feature"]
    value = updates / dt
    flops = searched * wl["flop"]
    kernel_s = st["kernel_ms"] / 1e3
    launches = st["kernel_launches"]
    achieved = flops / kernel_s / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": wl["traffic"],
            "kernel": wl["kernel"], "launches": launches,
            "mean_launch_us": 1e3 * st["kernel_ms"] / max(launches, 1),
            "flop_per_update": wl["flop"], "cbrt_per_update": wl["cbrt"]}
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(kind, xyz, cells, args.cpu_iters, threads, not kw["fixed_boundary"], args.workload)
    out = {
        "metric": METRIC, "value": value, "unit": "vertex updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "name": args.workload,
                   "vertex_updates_per_step": updates // max(args.steps, 1),
                   "searches": {"inner": st["searches_inner"], "surface": st["searches_surface"],
                                "feature": st["searches_feature"]},
                   "colors_inner": counts["n_colors_inner"], "colors_free": counts["n_colors_free"],
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
