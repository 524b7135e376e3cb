"""Benchmark: plain GPU DMO (vertex updates/s).

Default workload: the 40^3 x 6 tet cube (68,921 vertices, 384,000 tets,
interior jitter +-0.35h, boundary fixed, 8^3 candidates x 3 levels).  A step
is one DMO iteration over the whole mesh: every colour class searched once
(59,319 inner vertex searches for the tet cube).  Other workloads
(--workload): hex48 = 48^3 hex cube with the surface free (117,649
vertices), tri64 = 64 x 64 flat triangle grid (boundary fixed), surf_tri /
surf_quad = cube-sphere surface meshes of 96 x 96 cells per cube face
(55,298 vertices; triangles or quads) through the surface path.

The mesh is built and resident in HBM before the timed region; W warm-up
iterations run on a separate copy of the mesh, then K iterations of a fresh
mesh are timed (wall clock around the blocking iterate call, which returns
after the last kernel and the per-iteration quality check).

Output: one JSON line with the metric, a roofline object for the dominant
kernel (the inner search; canonical FP32 operations per vertex update x
updates per launch / mean launch duration from HIP events on the kernel
dispatches) and a CPU baseline (the oracle, a C/OpenMP restatement, on the
host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex updates/s at equal final min mean-ratio quality; learned vs plain DMO"
CPU_BUDGET_S = 10.0
# learned proposal window per element type: (levels, fraction of the plain
# first-level window), chosen by sweeps with the deferred fallback on the
# seed-7 and seed-8 40^3 tet / 48^3 hex cubes (the benchmark meshes use seed 0)
LEARN_WINDOW = {"tet": (1, 0.025), "hex": (1, 2.0)}
# proposal mode per element type (3: the network's window, then a quarter-size
# window around the vertex before the full search; 1: the network's window
# alone), chosen on the seed-7 / seed-8 cubes at K = 20: on the 40^3 tet cube
# mode 3 reaches the plain quality 1.3x faster than mode 1, on the 48^3 hex
# cube mode 1 is 1.2x faster than mode 3
LEARN_MODE = {"tet": 3, "hex": 1}
# the learned configuration as specified: one 4x4x4 grid on 1/4 of the plain
# first-level window, reported beside the tuned window
SPEC_WINDOW = (1, 0.25)
PEAK_FP32_TFLOPS = 157.3
# Work of one tet / hex vertex update in canonical FP32 operations: the
# candidates of a full search (3 levels x 512) and the start position, times
# the one-ring size (24 tets / 8 hexes on the structured cubes), times the
# operations the canonical evaluation executes per (candidate, ring element)
# with CPL candidates per lane (canon_flop); the cube root per candidate is
# counted separately.  Flat triangles: the reference formula's 23 FLOP per
# triangle evaluation, 196 ring evaluations x 6 triangles per update.
# HBM traffic per launch: rocprofv3 --pmc FETCH_SIZE x 1.99981 (the
# copy-kernel scale of bench.py --calibrate) + WRITE_SIZE x 1.0, mean over
# the dispatches of the default bench line's runs
# (profiles/learned_vs_plain_tet40_hex48.md; the fan kernels in
# profiles/k_search_fan2d_surf.md).
def canon_flop(kind: str, cand_per_lane: int) -> float:
    if kind == "tet":
        return 2 + 3 + 7 + 9.0 / cand_per_lane
    return 5 * (2 + 3 + 8 + 9.0 / cand_per_lane)


RING = {"tet": 24, "hex": 8}            # elements per interior vertex of the structured cubes
FULL_EVALS = 3 * 512 + 1                 # candidates of a full search + the start position
MLP_FLOP = 2 * (96 * 128 + 128 * 128 + 128 * 3)   # per network row (BF16 MFMA)
FLOP_TET = FULL_EVALS * RING["tet"] * canon_flop("tet", 8)
FLOP_HEX = FULL_EVALS * RING["hex"] * canon_flop("hex", 8)
# Surface-mesh searches (algorithmic FLOP per vertex update, valence n = 6
# triangles / 4 quads): per candidate the projection onto the blended local
# surfaces (57 + 69 n) plus the signed fan metric (59 n triangles, 104 n
# quads); 3 levels x 64 candidates (2D) or 3 x 8 (1D, projection 45 + 52 * 2
# curve neighbours).
SURF_FLOP = {"tri": 3 * 64 * (57 + 69 * 6 + 59 * 6), "quad": 3 * 64 * (57 + 69 * 4 + 104 * 4)}
# executed VALU issue rate of the search kernels (SQ_INSTS_VALU per launch /
# the kernel trace's mean duration / 256 CUs) over the highest rate measured
# for the search's own instruction mix (bench.py --calibrate's selection-step
# probe, 3.81 wave-instructions per CU per ns), from the counter passes of
# profiles/learned_vs_plain_tet40_hex48.md
VALU_ISSUE_FRAC = {"k_search_tet": 0.81, "k_search_hex": 0.93, "k_search_tet_learned_def": 0.37,
                   "k_search_hex_learned_def": 0.47}
VALU_ISSUE_BASIS = ("SQ_INSTS_VALU / kernel-trace duration / CUs over the calibrated mixed-instruction issue "
                    "ceiling (profiles/learned_vs_plain_tet40_hex48.md, profiles/valu_issue_calibration.md)")
SURF_FLOP1D = {"tri": 3 * 8 * (149 + 59 * 6), "quad": 3 * 8 * (149 + 104 * 4)}
WORKLOADS = {
    "tet40": {"kind": "tet", "flop": FLOP_TET, "cbrt": FULL_EVALS, "kernel": "k_search_tet<8>",
              "traffic": 9_346_689, "learned_traffic": 6_482_023},
    "hex48": {"kind": "hex", "flop": FLOP_HEX, "cbrt": FULL_EVALS, "kernel": "k_search_hex<8>",
              "traffic": 16_172_174, "learned_traffic": 10_517_734},
    # the 48^3 hex cube at jitter 0.15h: a valid mesh, so the surface and
    # feature passes run every iteration (the roofline is the inner kernel's)
    "hex48_surface": {"kind": "hex", "flop": FLOP_HEX, "cbrt": FULL_EVALS, "kernel": "k_search_hex<8>",
                      "traffic": None},
    "tri64": {"kind": "tri", "flop": 27_048, "cbrt": 0, "kernel": "k_search_flat<8>", "traffic": None},
    "surf_tri": {"kind": "tri", "surface": True, "flop": SURF_FLOP["tri"], "flop1d": SURF_FLOP1D["tri"], "cbrt": 0,
                 "kernel": "k_search_fan2d<false>", "traffic": 18_791_143},
    "surf_quad": {"kind": "quad", "surface": True, "flop": SURF_FLOP["quad"], "flop1d": SURF_FLOP1D["quad"], "cbrt": 0,
                  "kernel": "k_search_fan2d<true>", "traffic": 20_562_741},
    "batch256": {"kind": "batch", "flop": FLOP_TET, "cbrt": FULL_EVALS, "kernel": "k_batch_round", "traffic": None},
    # the default line: learned (+ batched) vs plain on tet40 and hex48
    "learned_vs_plain": {"kind": "tet", "flop": FLOP_TET, "cbrt": FULL_EVALS, "kernel": "k_search_tet_learned_def<8>",
                   "traffic": None},
}
# algorithmic FLOP of one flat-triangle vertex update (8x8 candidates: 196 ring
# evaluations x 6 triangles x 23 FLOP)
FLOP_PER_UPDATE_TRI = 27_048


def _batch_learned(args, b, dt_plain, q_plain, qm_plain):
    """The batch's tet meshes with the learned proposal (--batch-fallback policy,
    the tet window unless --learn-depth / --learn-frac) against the plain
    batch's K iterations, time to equal quality: K learned iterations from
    the start positions, then single iterations while any mesh is below the
    plain run's final minimum or mean quality (-1e-3), at most 2K in all.
    The triangle meshes keep the plain search, so the batch needs at least
    K iterations for them.  Network: trained on pairs recorded from plain
    DMO on the tet meshes of another mix (seed + 5000)."""
    from volumesurfacediscretemeshoptimization_amd import DMODataset, DMOMeshTetFactory, DMOMLP, DMOTetClass, synth
    t0 = time.perf_counter()
    ds = DMODataset(args.batch_pairs)
    used = 0
    for k, xyz, cells in synth.batch_mix(256, seed=args.seed + 5000):
        if k != "tet" or len(ds) >= args.batch_pairs:
            continue
        m = DMOMeshTetFactory.create(xyz, cells)
        o = DMOTetClass(m, n_iter=args.steps, fixed_boundary=True)
        o.record(ds, args.steps)
        o.close()
        m.close()
        used += 1
    rec_s = time.perf_counter() - t0
    mlp = DMOMLP(seed=args.seed + 1)
    t0 = time.perf_counter()
    losses = mlp.train(ds, epochs=args.batch_epochs, batch=4096, lr=1e-3, seed=args.seed + 2)
    train_s = time.perf_counter() - t0
    n_pairs = len(ds)
    ds.close()
    depth, frac = args.learn_depth, args.learn_frac
    b.restore()
    b.set_fallback(args.batch_fallback)
    b.set_proposal(mlp, mode=args.learn_mode, depth=depth, window_frac=frac)
    b.iterate(max(args.warmup, 1))           # warm-up, then back to the start positions
    b.restore()
    ls0 = b.learned_stats()
    st0 = b.stats()
    t0 = time.perf_counter()
    b.iterate(args.steps)
    q_l, qm_l = b.min_quality(), b.mean_quality()
    below_k = int(np.sum((q_l < q_plain - 1e-3) | (qm_l < qm_plain - 1e-3)))
    below, its = below_k, args.steps
    # meshes still below the plain run's final quality: further iterations
    # (timed, with their checks) until every mesh reaches it, at most K more
    while below > 0 and its < 2 * args.steps:
        b.iterate(1)
        its += 1
        q_l, qm_l = b.min_quality(), b.mean_quality()
        below = int(np.sum((q_l < q_plain - 1e-3) | (qm_l < qm_plain - 1e-3)))
    dt = time.perf_counter() - t0
    st = b.stats()
    ls = b.learned_stats()
    b.set_proposal(None, mode=0)
    mlp.close()
    visits = ls["learned_visits"] - ls0["learned_visits"]
    return {"policy": args.batch_fallback, "levels": depth, "window_frac": frac, "plain_iterations": args.steps,
            "learned_iterations": its, "meshes_below_plain_quality_at_plain_iterations": below_k,
            "learned_ms_per_step": 1e3 * dt / its, "plain_ms_per_step": 1e3 * dt_plain / args.steps,
            "speedup_vs_plain": dt_plain / dt if below == 0 else 0.0,
            "meshes_below_plain_quality": below,
            "min_quality_end_worst": float(q_l.min()), "device_ms_per_step": (st["gpu_ms"] - st0["gpu_ms"]) / its,
            "learned_visits_per_step": visits // its,
            "fallback_rate": (ls["fallbacks"] - ls0["fallbacks"]) / max(1, visits),
            "training_pairs": n_pairs, "training_meshes": used, "record_s": rec_s, "train_s": train_s,
            "loss_last_epoch": float(losses[-1]) if len(losses) else None}


def bench_batch(args):
    """256 meshes (half flat triangle grids, half tet cubes, vertex counts
    log-uniform in [500, 50000]) swept colour round by colour round, one
    iteration replayed as a hipGraph, parameters from the autotuner (run
    before the timed region).  The unbatched baseline (one optimiser per
    mesh, meshes one after another) is timed beside it."""
    from volumesurfacediscretemeshoptimization_amd import (DMOBatch, DMOMeshTetFactory, DMOMeshTriFactory,
                                                           DMOTetClass, DMOTriFlatClass, synth)
    mix = synth.batch_mix(256, seed=args.seed)
    fac = {"tri": DMOMeshTriFactory, "tet": DMOMeshTetFactory}
    t0 = time.perf_counter()
    meshes = [fac[k].create(xyz, cells) for k, xyz, cells in mix]
    build_s = time.perf_counter() - t0
    b = DMOBatch(meshes, use_graph=bool(args.batch_graph))
    q0 = b.min_quality()
    if args.tune_iters > 0:
        chosen = b.autotune(iters=args.tune_iters, tol=1e-3, log_path=args.tune_log)
    else:
        # the configuration the autotuner keeps on this mix (nq 8, depth 3, one
        # wave per vertex, one block per item), without its trial graphs
        b.set_params(nq=8, depth=3, waves_per_vertex=1, persistent=0)
        chosen = b.params()
    # warm-up: one iteration, then back to the start positions
    b.iterate(max(args.warmup, 1))
    b.restore()
    st0 = b.stats()
    t0 = time.perf_counter()
    b.iterate(args.steps)
    dt = time.perf_counter() - t0
    st = b.stats()
    q1 = b.min_quality()
    qm1 = b.mean_quality()
    searches = st["searches"] - st0["searches"]
    gpu_s = (st["gpu_ms"] - st0["gpu_ms"]) / 1e3
    # work per iteration: tet searches x canonical tet operations + tri searches x tri FLOP
    per_mesh = []
    for i, (k, _, _) in enumerate(mix):
        c = meshes[i].counts()
        per_mesh.append(c["n_inner"] if k == "tet" else c["n_free"])
    tet_s = sum(n for n, (k, _, _) in zip(per_mesh, mix) if k == "tet")
    tri_s = sum(n for n, (k, _, _) in zip(per_mesh, mix) if k == "tri")
    flops = args.steps * (tet_s * WORKLOADS["tet40"]["flop"] + tri_s * FLOP_PER_UPDATE_TRI)
    achieved = flops / gpu_s / 1e12
    # unbatched baseline with the reference parameters (8, 3)
    for k, xyz, cells in mix[:2]:          # warm the single-mesh kernels on throwaway copies
        wm = fac[k].create(xyz, cells)
        wo = (DMOTriFlatClass if k == "tri" else DMOTetClass)(wm, n_iter=1, fixed_boundary=True)
        wo.iterate(1)
        wo.close()
        wm.close()
    plain_meshes = [fac[k].create(xyz, cells) for k, xyz, cells in mix]
    opts = [(DMOTriFlatClass if k == "tri" else DMOTetClass)(m, n_iter=args.steps, fixed_boundary=True)
            for (k, _, _), m in zip(mix, plain_meshes)]
    t1 = time.perf_counter()
    for o in opts:
        o.iterate(args.steps)
    dt_plain = time.perf_counter() - t1
    plain_updates = sum(o.stats()["searches_inner"] + o.stats()["searches_surface"] for o in opts)
    q_plain = np.array([o.getQuality() for o in opts], dtype=np.float32)
    learned = None
    if args.batch_learned:
        try:
            learned = _batch_learned(args, b, dt, q1, qm1)
        except Exception as e:        # the plain line must not depend on it
            learned = {"error": f"{type(e).__name__}: {e}"}
    cpu = None
    if not args.no_cpu:
        threads = host_cpu()["threads_all"]
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py
        n = 0
        tc = time.perf_counter()
        sample = 0
        for k, xyz, cells in mix:
            om = oracle_py.OracleMesh(k, xyz, cells)
            n += om.iterate(nthreads=threads)
            sample += 1
            if time.perf_counter() - tc > 15.0:
                break
        dtc = time.perf_counter() - tc
        cpu = {"value": n / dtc, "unit": "vertex updates/s", "cores": threads, "kind": "port",
               "sample": f"one DMO iteration of each of the first {sample} meshes of the mix ({n} vertex searches, "
                         f"{dtc:.1f} s) by the C/OpenMP oracle with {threads} threads"}
    out = {
        "metric": METRIC, "value": searches / dt, "unit": "vertex updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched DMO, 256 meshes (even: flat triangle grids, odd: tet cubes; vertex counts "
                               f"log-uniform in [500, 50000], seed {args.seed}; jitter +-0.35h), boundary fixed, one "
                               "persistent launch per colour round, iteration replayed as a hipGraph, autotuned",
                   "name": "batch256", "meshes": st["meshes"], "total_vertices": st["total_vertices"],
                   "rounds": st["rounds"], "vertex_updates_per_step": st["items_per_iteration"],
                   "autotuned": {k: chosen[k] for k in ("nq", "depth", "waves_per_vertex", "persistent")},
                   "mesh_build_s": build_s, "min_quality_start_worst": float(q0.min()),
                   "min_quality_end_worst": float(q1.min()),
                   "unbatched_value": plain_updates / dt_plain, "unbatched_ms_per_step": 1e3 * dt_plain / args.steps,
                   "unbatched_min_quality_end_worst": float(q_plain.min()),
                   "batched_over_unbatched": (searches / dt) / (plain_updates / dt_plain),
                   "learned_tets_vs_plain": learned,
                   "tune_log": args.tune_log},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "k_batch_typed (all rounds + "
                     "per-mesh quality, device time of the replayed iterations)",
                     "flop_basis": "canonical FP32 operations per tet update (13.125 per candidate and ring "
                                   "element), 23 FLOP per flat-triangle evaluation"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    b.close()
    for o in opts:
        o.close()
    return 0


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
    if name == "hex48_surface":
        xyz, cells = synth.hex_cube(48, seed=seed, jitter=0.15)
        desc = ("plain GPU DMO, hex cube 48^3 cells (117,649 vertices, 110,592 hexes), interior jitter +-0.15h, "
                f"surface vertices +-0.15h tangentially, seed {seed}, surface free: a valid mesh, so the surface "
                "(quadric fits, 8 x 8 candidates x 3 levels, projection) and feature (8 x 3) passes run every "
                "iteration after the inner passes; one step = one iteration")
        return xyz, cells, desc, {"fixed_boundary": False}
    if name in ("surf_tri", "surf_quad"):
        quads = name == "surf_quad"
        xyz, cells = synth.cube_surface(96, seed=seed, jitter=0.3, quads=quads, sphere=True)
        desc = (f"plain GPU DMO on a surface mesh: cube-sphere, 96 x 96 {'quads' if quads else 'quads split in two triangles'} "
                f"per cube face ({len(xyz):,} vertices, {len(cells):,} {'quads' if quads else 'triangles'}), face "
                f"vertices jittered +-0.3h tangentially before the radial projection, seed {seed}; free vertices "
                "search 8 x 8 candidates x 3 levels on the blended local quadrics, feature vertices 8 x 3 on their "
                "curves; one step = one iteration (all colour classes, nearest-neighbour walk, normal update)")
        return xyz, cells, desc, {"fixed_boundary": False}
    if name == "tri64":
        xyz, cells = synth.tri_grid(64, seed=seed)
        desc = ("plain GPU DMO, flat triangle grid 64 x 64 vertices (7,938 triangles), interior jitter +-0.4h, "
                f"seed {seed}, boundary fixed; 8 x 8 candidates x 3 levels; one step = one iteration")
        return xyz, cells, desc, {"fixed_boundary": True}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(kind, xyz, cells, iters: int, threads: int, surface: bool, name: str, surface_mesh: bool = False):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    m = oracle_py.OracleMesh(kind, xyz, cells)
    if surface or surface_mesh:
        m.surface_init()
    t0 = time.perf_counter()
    n = 0
    done = 0
    # whole iterations until the CPU time budget (iters > 0: exactly iters)
    while (done < iters) if iters > 0 else (done == 0 or time.perf_counter() - t0 < CPU_BUDGET_S):
        done += 1
        if surface_mesh:
            n2, n1 = m.iterate_surface_mesh(nthreads=threads)
            n += n2 + n1
        elif surface:
            n2, n1 = m.iterate_surface(nthreads=threads)
            n += m.m.nInner + n2 + n1
        else:
            n += m.iterate(nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "vertex updates/s", "cores": threads, "kind": "port",
            "sample": f"{done} full DMO iteration(s) of the same {name} mesh ({n} vertex searches, {dt:.1f} s) "
                      f"by the C/OpenMP oracle with {threads} threads"}


def learned_comparison(args):
    """Learned proposal vs plain DMO on the bench mesh; returns a dict (see
    bench_learned).  Never raises: an error is reported in the dict."""
    try:
        return _learned(args)
    except Exception as e:          # the plain line must not depend on it
        return {"error": f"{type(e).__name__}: {e}"}


def bench_learned(args):
    """Learned proposal vs plain DMO on the same mesh.  Untimed setup: training
    pairs recorded from plain DMO on seeded meshes of the same family (seeds
    1000+), then Adam training of the MLP.  Timed: the plain run of K
    iterations, then the learned run until its minimum and mean quality reach
    the plain run's final ones (-1e-3).  value = plain-equivalent vertex
    updates per second at equal final quality (K x inner vertices / learned
    wall time)."""
    r = _learned(args)
    xyz, cells, desc, kw = make_mesh(args.workload, args.seed)
    wl = WORKLOADS[args.workload]
    kind = wl["kind"]
    out = {
        "metric": METRIC, "value": r["value"], "unit": "vertex updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["learned_ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 search, bf16 MLP", "data": "synthetic",
        "config": dict(workload=desc + f"; learned proposal ({args.learn_depth} levels of 4x4x4 on "
                       f"{args.learn_frac} of the plain first-level window, {args.learn_fallback} fallback to the "
                       "full search) vs plain", name=args.workload,
                       mode="learned", **{k: v for k, v in r.items() if k not in ("value", "roofline")}),
        "roofline": r["roofline"],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = host_cpu()["threads_all"]
        out["cpu_baseline"] = cpu_baseline(kind, xyz, cells, args.cpu_iters, threads, not kw["fixed_boundary"],
                                           args.workload)
    print(json.dumps(out), flush=True)
    return 0


def _learned_batched(args, mlp, fac, xyz, cells, kind, single_plain_s):
    """Learned + batched on the bench mesh: a one-mesh DMOBatch (one
    iteration replayed as a hipGraph, no host check between iterations),
    the learned proposal with the line's fallback policy against the same
    batch's plain search.  The learned iteration count n is the first at
    which the min and mean quality reach the plain run's final ones (-1e-3),
    found in an untimed run (the runs are deterministic, bit for bit); the
    timed runs are one replay call each, K plain iterations and n learned
    ones, and the timed learned run's final quality is checked again.  The
    ratio to the single-mesh optimiser's plain K iterations (single_plain_s)
    is reported beside it."""
    from volumesurfacediscretemeshoptimization_amd import DMOBatch, _lib
    elem = _lib.VSDMO_TET if kind == "tet" else _lib.VSDMO_HEX
    m = fac.create(xyz, cells)
    # one block per item (one launch per element type): the schedule the
    # autotuner keeps on the 256-mesh batch
    b = DMOBatch([m], use_graph=bool(args.batched_graph), persistent=False)
    b.iterate(1)                             # warm-up (captures the graph)
    b.restore()
    t0 = time.perf_counter()
    b.iterate(args.steps)
    dt_plain = time.perf_counter() - t0
    qp, qmp = float(b.min_quality()[0]), float(b.mean_quality()[0])
    b.restore()
    b.set_fallback(args.learn_fallback)
    b.set_proposal(mlp, mode=args.learn_mode, depth=args.learn_depth, window_frac=args.learn_frac, elem=elem)
    n = 0
    while n < 3 * args.steps:
        b.iterate(1)
        n += 1
        if b.min_quality()[0] >= qp - 1e-3 and b.mean_quality()[0] >= qmp - 1e-3:
            break
    b.restore()
    ls0 = b.learned_stats()
    t0 = time.perf_counter()
    b.iterate(n)
    dt = time.perf_counter() - t0
    ql, qml = float(b.min_quality()[0]), float(b.mean_quality()[0])
    ls = b.learned_stats()
    reached = ql >= qp - 1e-3 and qml >= qmp - 1e-3
    b.set_proposal(None, mode=0, elem=elem)
    b.close()
    m.close()
    visits = ls["learned_visits"] - ls0["learned_visits"]
    return {"policy": args.learn_fallback, "mode": args.learn_mode, "plain_iterations": args.steps,
            "learned_iterations": n,
            "plain_ms_per_step": 1e3 * dt_plain / args.steps, "learned_ms_per_step": 1e3 * dt / n,
            "speedup_vs_plain": dt_plain / dt if reached else 0.0,
            "speedup_vs_single_mesh_plain": single_plain_s / dt if reached else 0.0,
            "reached_plain_quality": bool(reached),
            "min_quality_plain": qp, "min_quality_learned": ql, "mean_quality_plain": qmp,
            "mean_quality_learned": qml, "fallback_rate": (ls["fallbacks"] - ls0["fallbacks"]) / max(1, visits)}


def _learned(args):
    from volumesurfacediscretemeshoptimization_amd import (DMODataset, DMOHexClass, DMOMeshHexFactory,
                                                           DMOMeshTetFactory, DMOMLP, DMOTetClass)
    kind = WORKLOADS[args.workload]["kind"]
    fac, cls = (DMOMeshTetFactory, DMOTetClass) if kind == "tet" else (DMOMeshHexFactory, DMOHexClass)
    xyz, cells, desc, kw = make_mesh(args.workload, args.seed)
    kw = {"fixed_boundary": kw["fixed_boundary"], "waves_per_vertex": args.waves}
    # training pairs
    t0 = time.perf_counter()
    ds = DMODataset(args.pairs)
    s = 0
    while len(ds) < args.pairs:
        txyz, tcells, _, _ = make_mesh(args.workload, 1000 + s)
        m = fac.create(txyz, tcells)
        o = cls(m, n_iter=args.steps, **kw)
        o.record(ds, args.steps)
        o.close()
        m.close()
        s += 1
    rec_s = time.perf_counter() - t0
    mlp = DMOMLP(seed=args.seed + 1)
    t0 = time.perf_counter()
    losses = mlp.train(ds, epochs=args.epochs, batch=4096, lr=1e-3, seed=args.seed + 2)
    train_s = time.perf_counter() - t0
    n_pairs = len(ds)
    ds.close()
    # warm-up of both paths (and both fallback policies) on a throwaway copy
    wm = fac.create(xyz, cells)
    wo = cls(wm, n_iter=4, **kw)
    for pol in ("immediate", "deferred"):
        wo.set_proposal(mlp, mode=args.learn_mode, depth=args.learn_depth, window_frac=args.learn_frac, fallback=pol)
        wo.iterate(1)
    wo.set_proposal(None, mode=0)
    wo.iterate(1)
    wo.close()
    wm.close()
    # plain
    mp = fac.create(xyz, cells)
    op = cls(mp, n_iter=args.steps, **kw)
    q0 = op.getQuality()
    t0 = time.perf_counter()
    op.iterate(args.steps)
    dt_plain = time.perf_counter() - t0
    stp = op.stats()
    q_plain, qm_plain = op.getQuality(), op.getMeanQuality()
    plain_updates = stp["searches_inner"] + stp["searches_surface"] + stp["searches_feature"]
    # learned (wall clock without per-launch timing events, like the plain run)
    def to_quality(policy, depth=args.learn_depth, frac=args.learn_frac):
        ml = fac.create(xyz, cells)
        ol = cls(ml, n_iter=3 * args.steps, **kw)
        ol.set_proposal(mlp, mode=args.learn_mode, depth=depth, window_frac=frac, fallback=policy)
        its = 0
        t0 = time.perf_counter()
        while its < 3 * args.steps:
            ol.doIteration()
            its += 1
            # equal quality: minimum and mean within 1e-3 of the plain run's
            if ol.getQuality() >= q_plain - 1e-3 and ol.getMeanQuality() >= qm_plain - 1e-3:
                break
        dt = time.perf_counter() - t0
        out = (its, dt, ol.stats(), ol.getQuality(), ol.getMeanQuality())
        ol.close()
        ml.close()
        return out

    its, dt_learned, stl, q_l, qm_l = to_quality(args.learn_fallback)
    other = "immediate" if args.learn_fallback == "deferred" else "deferred"
    extras = {}
    if args.learned_extras:
        for key, (pol, depth, frac) in ((f"{other}_fallback", (other, args.learn_depth, args.learn_frac)),
                                        ("specified_window", (args.learn_fallback,) + SPEC_WINDOW)):
            its_x, dt_x, st_x, q_x, qm_x = to_quality(pol, depth, frac)
            ok_x = q_x >= q_plain - 1e-3 and qm_x >= qm_plain - 1e-3
            extras[key] = {"fallback": pol, "levels": depth, "window_frac": frac, "learned_iterations": its_x,
                           "learned_ms_per_step": 1e3 * dt_x / its_x, "reached_plain_quality": bool(ok_x),
                           "speedup_vs_plain": (plain_updates / dt_x) / (plain_updates / dt_plain) if ok_x else 0.0,
                           "fallback_rate": st_x["fallbacks"] / max(1, st_x["searches_learned"])}
    # the same run again with HIP events around every launch, for the kernel time
    mk = fac.create(xyz, cells)
    ok = cls(mk, n_iter=its, **kw)
    ok.set_proposal(mlp, mode=args.learn_mode, depth=args.learn_depth, window_frac=args.learn_frac, fallback=args.learn_fallback)
    ok.set_timing(True)
    ok.iterate(its)
    stk = ok.stats()
    ok.close()
    mk.close()
    batched = None
    if args.learned_batched:
        try:
            batched = _learned_batched(args, mlp, fac, xyz, cells, kind, dt_plain)
        except Exception as e:    # the learned line must not depend on it
            batched = {"error": f"{type(e).__name__}: {e}"}
    learned_updates = stl["searches_inner"] + stl["searches_surface"] + stl["searches_feature"]
    reached = q_l >= q_plain - 1e-3 and qm_l >= qm_plain - 1e-3
    value = plain_updates / dt_learned if reached else 0.0
    # canonical FP32 operations of the windows (4 candidates per lane) and of
    # the full searches of the fallbacks (8 per lane)
    ring = RING[kind]
    flops = (stl["searches_learned"] * (64 * args.learn_depth + 1) * ring * canon_flop(kind, 4)
             + stl["fallbacks"] * FULL_EVALS * ring * canon_flop(kind, 8))
    achieved = flops / (stk["kernel_ms"] / 1e3) / 1e12
    op.close()
    mp.close()
    mlp.close()
    return {
        "value": value,
        "value_definition": "plain run's vertex updates / learned wall time to reach the plain run's final min and "
                            "mean quality (-1e-3)",
        "plain_value": plain_updates / dt_plain, "plain_ms_per_step": 1e3 * dt_plain / args.steps,
        "learned_ms_per_step": 1e3 * dt_learned / its,
        "plain_iterations": args.steps, "learned_iterations": its,
        "learned_raw_value": learned_updates / dt_learned,
        "speedup_vs_plain": (value / (plain_updates / dt_plain)) if reached else 0.0,
        "reached_plain_quality": bool(reached), "min_quality_start": q0,
        "min_quality_plain": q_plain, "min_quality_learned": q_l,
        "mean_quality_plain": qm_plain, "mean_quality_learned": qm_l,
        "fallback_policy": args.learn_fallback,
        "fallback_rate": stl["fallbacks"] / max(1, stl["searches_learned"]),
        **extras,
        "learned_batched": batched,
        "training_pairs": n_pairs, "training_meshes": s, "record_s": rec_s, "train_s": train_s,
        "epochs": args.epochs, "loss_first_epoch": float(losses[0]) if len(losses) else None,
        "loss_last_epoch": float(losses[-1]) if len(losses) else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": (WORKLOADS[args.workload].get("learned_traffic")
                                 if args.learn_fallback == "deferred" else None),
                     "kernel": (f"k_search_{kind}_learned_def<8> (network, proposal windows and the deferred "
                                "full searches in one launch per class)" if args.learn_fallback == "deferred" else
                                f"k_search_{kind}_learned<8> + k_search_list_{kind}<8, 4> (windows, then the "
                                "class's full searches)"),
                     "launches": stk["kernel_launches"],
                     "mean_launch_us": 1e3 * stk["kernel_ms"] / max(1, stk["kernel_launches"])},
    }


# ---------------------------------------------------------------------------
# north-star line: learned (+ batched) vs plain DMO at equal final quality on
# the 40^3 tet cube and the 48^3 hex cube
# ---------------------------------------------------------------------------
# FP32 operations of the canonical evaluation per (candidate, ring element),
# counted from the kernels' inner loops (add / sub / mul / compare / min = 1,
# fma = 2; selects and moves not counted).  Tet: det = fma (2), L = add + fma
# (3), the running selection d^2, L^3, two cross products, compare, min (7),
# plus the lane's (y, z) terms of the element (two fma for det, add + two fma
# for L: 9) shared by its candidates per lane (8 in the full search, 4 in the
# proposal window).  Hex, per p-dependent sub-tet (5 per hex): det (2), F
# (3), selection with both minima (8), plus 9 shared terms.
def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    share = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    return {"model": model, "nproc": os.cpu_count(), "affinity": avail, "threads_all": share}


def cpu_baseline_tet(xyz, cells, name: str, reps: int = 3):
    """The oracle (C/OpenMP restatement of the reference search) on the host
    cores: whole iterations with every core of the job's share, and with one
    core a sample of each iteration (the first 1/8 of every colour class; the
    structured cube's vertices all cost the same), median of `reps` runs each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    cpu = host_cpu()
    res = {}
    for label, threads in (("all", cpu["threads_all"]), ("one", 1)):
        om = oracle_py.OracleMesh("tet", xyz, cells)
        off = om.offsets("inner")
        cap = 0 if label == "all" else int(np.max(np.diff(off))) // 8 + 1
        rates, secs = [], 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            n = om.iterate(nthreads=threads, max_per_class=cap)
            dt = time.perf_counter() - t0
            secs += dt
            rates.append(n / dt)
        res[label] = {"threads": threads, "rate": float(np.median(rates)), "rates": rates, "seconds": secs,
                      "per_class_cap": cap}
    return {"value": res["all"]["rate"], "unit": "vertex updates/s", "cores": res["all"]["threads"], "kind": "port",
            "value_one_core": res["one"]["rate"],
            "sample": (f"{name}: {reps} whole DMO iterations with {res['all']['threads']} threads (median "
                       f"{res['all']['rate']:.0f}/s, {res['all']['seconds']:.1f} s), and {reps} iterations on 1 "
                       f"thread over the first {res['one']['per_class_cap']} vertices of each colour class "
                       f"(median {res['one']['rate']:.0f}/s, {res['one']['seconds']:.1f} s), C/OpenMP oracle"),
            "cpu_model": cpu["model"], "nproc": cpu["nproc"], "affinity_cpus": cpu["affinity"],
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _train_net(name, args, fac, cls, kw, pairs, epochs):
    """pairs recorded from plain DMO (K iterations each) on meshes of the same
    family with seeds 1000, 1001, ..., then Adam training on the GPU"""
    from volumesurfacediscretemeshoptimization_amd import DMODataset, DMOMLP
    t0 = time.perf_counter()
    ds = DMODataset(pairs)
    s = 0
    while len(ds) < pairs:
        txyz, tcells, _, _ = make_mesh(name, 1000 + s)
        m = fac.create(txyz, tcells)
        o = cls(m, n_iter=args.steps, **kw)
        o.record(ds, args.steps)
        o.close()
        m.close()
        s += 1
    rec_s = time.perf_counter() - t0
    mlp = DMOMLP(seed=args.seed + 1)
    t0 = time.perf_counter()
    losses = mlp.train(ds, epochs=epochs, batch=4096, lr=1e-3, seed=args.seed + 2)
    info = {"training_pairs": len(ds), "training_meshes": s, "record_s": rec_s, "train_s": time.perf_counter() - t0,
            "epochs": epochs, "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1])}
    ds.close()
    return mlp, info


def _to_quality(fac, cls, xyz, cells, kw, mlp, mode, depth, frac, policy, qp, qmp, max_its, timing=False):
    """iterations of one optimiser with the proposal until its minimum and
    mean quality reach qp, qmp (-1e-3) or max_its; wall time of the loop"""
    m = fac.create(xyz, cells)
    o = cls(m, n_iter=max_its, **kw)
    o.set_proposal(mlp if mode in (1, 3) else None, mode=mode, depth=depth, window_frac=frac, fallback=policy)
    if timing:
        o.set_timing(True)
    its, hist = 0, []
    t0 = time.perf_counter()
    while its < max_its:
        o.doIteration()
        its += 1
        if timing:
            st = o.stats()
            hist.append({"min": o.getQuality(), "mean": o.getMeanQuality(), "fallbacks": st["fallbacks"],
                         "visits": st["searches_learned"], "kernel_ms": st["kernel_ms"]})
        if o.getQuality() >= qp - 1e-3 and o.getMeanQuality() >= qmp - 1e-3:
            break
    dt = time.perf_counter() - t0
    st = o.stats()
    out = {"iterations": its, "seconds": dt, "min_quality": o.getQuality(), "mean_quality": o.getMeanQuality(),
           "reached": o.getQuality() >= qp - 1e-3 and o.getMeanQuality() >= qmp - 1e-3,
           "visits": st["searches_learned"], "fallbacks": st["fallbacks"], "kernel_ms": st["kernel_ms"],
           "launches": st["kernel_launches"], "history": hist}
    o.close()
    m.close()
    return out


def learned_vs_plain_mesh(name, args):
    """plain vs learned vs learned + batched on one mesh at equal final quality"""
    from volumesurfacediscretemeshoptimization_amd import (DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                           DMOTetClass)
    kind = WORKLOADS[name]["kind"]
    fac, cls = (DMOMeshTetFactory, DMOTetClass) if kind == "tet" else (DMOMeshHexFactory, DMOHexClass)
    xyz, cells, desc, kw = make_mesh(name, args.seed)
    kw = {"fixed_boundary": kw["fixed_boundary"]}
    depth, frac = LEARN_WINDOW[kind]
    mode = args.learn_mode or LEARN_MODE[kind]
    policy = args.learn_fallback
    K = args.steps
    mlp, train = _train_net(name, args, fac, cls, kw, args.pairs, args.epochs)
    # warm-up: W iterations of each path on throwaway copies
    for learned in (False, True):
        wm = fac.create(xyz, cells)
        wo = cls(wm, n_iter=max(args.warmup, 1), **kw)
        if learned:
            wo.set_proposal(mlp, mode=mode, depth=depth, window_frac=frac, fallback=policy)
        wo.iterate(max(args.warmup, 1))
        wo.close()
        wm.close()
    # plain, timed (no per-launch events), then its kernel time with events
    tb = time.perf_counter()
    mp = fac.create(xyz, cells)
    build_s = time.perf_counter() - tb
    op = cls(mp, n_iter=K, **kw)
    q0 = op.getQuality()
    t0 = time.perf_counter()
    op.iterate(K)
    dt_plain = time.perf_counter() - t0
    stp = op.stats()
    qp, qmp = op.getQuality(), op.getMeanQuality()
    plain_updates = stp["searches_inner"] + stp["searches_surface"] + stp["searches_feature"]
    op.close()
    mp.close()
    mk = fac.create(xyz, cells)
    ok = cls(mk, n_iter=K, **kw)
    ok.set_timing(True)
    ok.iterate(K)
    stk = ok.stats()
    ok.close()
    mk.close()
    plain_kernel_us = 1e3 * stk["kernel_ms"] / max(1, stk["kernel_launches"])
    # learned, timed; the same run with events for the kernel time and the
    # per-iteration fallbacks; the headline mode and the other network mode
    other = 1 if mode == 3 else 3
    lrn = _to_quality(fac, cls, xyz, cells, kw, mlp, mode, depth, frac, policy, qp, qmp, 3 * K)
    lrk = _to_quality(fac, cls, xyz, cells, kw, mlp, mode, depth, frac, policy, qp, qmp, lrn["iterations"],
                      timing=True)
    lro = _to_quality(fac, cls, xyz, cells, kw, mlp, other, depth, frac, policy, qp, qmp, 3 * K)
    # the same windowed search around the vertex itself (no network)
    abl = _to_quality(fac, cls, xyz, cells, kw, None, 2, depth, frac, policy, qp, qmp, 3 * K)
    # the specified configuration: one level on 1/4 of the plain window
    spec = _to_quality(fac, cls, xyz, cells, kw, mlp, mode, SPEC_WINDOW[0], SPEC_WINDOW[1], policy, qp, qmp, 3 * K)
    # learned + batched: one-mesh batch, one iteration replayed as a graph
    saved = (args.learn_depth, args.learn_frac, args.learn_mode)
    args.learn_depth, args.learn_frac, args.learn_mode = depth, frac, mode
    try:
        lb = _learned_batched(args, mlp, fac, xyz, cells, kind, dt_plain)
    except Exception as e:          # reported, the other figures stand
        lb = {"error": f"{type(e).__name__}: {e}"}
    args.learn_depth, args.learn_frac, args.learn_mode = saved
    mlp.close()

    def rate(r):
        return plain_updates / r["seconds"] if r["reached"] else 0.0

    # the learned class kernel: canonical FP32 work of its windows and of the
    # full searches it runs for fallbacks, per launch
    ring = RING[kind]
    win_evals = lrk["visits"] * (64 * depth + 1)
    flop = (win_evals * ring * canon_flop(kind, 4) + lrk["fallbacks"] * FULL_EVALS * ring * canon_flop(kind, 8))
    launches = max(1, lrk["launches"])
    learned_kernel_us = 1e3 * lrk["kernel_ms"] / launches
    wl = WORKLOADS[name]
    plain_flop_launch = plain_updates / max(1, stk["kernel_launches"]) * FULL_EVALS * ring * canon_flop(kind, 8)
    res = {
        "workload": desc,
        "plain": {"iterations": K, "ms_per_step": 1e3 * dt_plain / K, "value": plain_updates / dt_plain,
                  "min_quality_start": q0, "min_quality": qp, "mean_quality": qmp,
                  "kernel": wl["kernel"], "kernel_mean_launch_us": plain_kernel_us,
                  "kernel_tflops_canonical": plain_flop_launch / (plain_kernel_us * 1e-6) / 1e12,
                  "mesh_build_ms": 1e3 * build_s},
        "learned": {"mode": mode, "window": [depth, frac], "fallback_policy": policy,
                    "iterations": lrn["iterations"],
                    "ms_per_learned_iteration": 1e3 * lrn["seconds"] / lrn["iterations"],
                    "value": rate(lrn), "speedup_vs_plain": dt_plain / lrn["seconds"] if lrn["reached"] else 0.0,
                    "reached_plain_quality": lrn["reached"], "min_quality": lrn["min_quality"],
                    "mean_quality": lrn["mean_quality"],
                    "fallback_rate": lrn["fallbacks"] / max(1, lrn["visits"]),
                    "kernel": f"k_search_{kind}_learned_def<8,true>",
                    "kernel_mean_launch_us": learned_kernel_us, "kernel_launches": lrk["launches"],
                    "kernel_tflops_canonical": flop / (lrk["kernel_ms"] * 1e-3) / 1e12,
                    "mfma_tflops": lrk["visits"] * MLP_FLOP / (lrk["kernel_ms"] * 1e-3) / 1e12,
                    "per_iteration": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in h.items()}
                                      for h in lrk["history"]]},
        "learned_batched": lb,
        f"learned_mode{other}": {"iterations": lro["iterations"], "reached_plain_quality": lro["reached"],
                                 "speedup_vs_plain": dt_plain / lro["seconds"] if lro["reached"] else 0.0,
                                 "ms_per_learned_iteration": 1e3 * lro["seconds"] / lro["iterations"],
                                 "fallback_rate": lro["fallbacks"] / max(1, lro["visits"])},
        "ablation_same_window_no_network": {
            "iterations": abl["iterations"], "reached_plain_quality": abl["reached"],
            "speedup_vs_plain": dt_plain / abl["seconds"] if abl["reached"] else 0.0,
            "min_quality": abl["min_quality"], "mean_quality": abl["mean_quality"],
            "fallback_rate": abl["fallbacks"] / max(1, abl["visits"]),
            "network_gain": (abl["seconds"] / lrn["seconds"]) if (abl["reached"] and lrn["reached"]) else None},
        "specified_window": {"window": list(SPEC_WINDOW), "iterations": spec["iterations"],
                             "reached_plain_quality": spec["reached"],
                             "speedup_vs_plain": dt_plain / spec["seconds"] if spec["reached"] else 0.0,
                             "fallback_rate": spec["fallbacks"] / max(1, spec["visits"])},
        "network": train,
        "_plain_updates": plain_updates, "_dt_plain": dt_plain, "_learned_flop_launch": flop / launches,
        "_learned_kernel_us": learned_kernel_us, "_xyz": xyz, "_cells": cells,
    }
    if kind == "hex":
        res["note"] = ("the jitter leaves inverted hexes that no search repairs: the minimum quality is 0 in every "
                       "run, so equal quality is decided by the mean")
    return res


CALIB_LIB = os.path.join(ROOT, "bench_calibrate.so")


PROBES = ("v_fma_f32", "v_mul_f32", "v_add_f32", "v_cmp_lt_f32+v_cndmask_b32", "tet selection step (mixed)")


def calibrate(copy_mib: int = 256, copy_reps: int = 5, iters: int = 20000, reps: int = 5):
    """(1) a device-to-device copy of copy_mib MiB (copy_reps kernels), whose
    HBM traffic is exactly known, to scale the rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings of the DMO kernels; (2) bench_calibrate.hip's
    k_valu_probe kernels, each a known count of one non-packed VALU
    instruction class on every lane (FP32 FMA at 2, 4 and 8 waves per SIMD,
    the others at 8), timed with HIP events: the VALU issue rates the search
    kernels are compared with.  Prints one JSON line."""
    import ctypes
    import torch
    n = copy_mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(copy_reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    if not os.path.exists(CALIB_LIB):
        raise SystemExit(f"{CALIB_LIB} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(CALIB_LIB)
    lib.calib_probe_run.restype = ctypes.c_int
    lib.calib_probe_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_float)]
    lib.calib_fma_per_lane.restype = ctypes.c_longlong
    lib.calib_fma_per_lane.argtypes = [ctypes.c_int]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    out = {"copy": {"mib": copy_mib, "kernels": copy_reps}, "cus": cus, "probes": []}
    runs = [(0, w) for w in (2, 4, 8)] + [(op, 8) for op in (1, 2, 3, 4)]
    for op, waves_per_simd in runs:
        blocks = cus * waves_per_simd           # 256-thread blocks: 4 waves, one per SIMD
        ms = (ctypes.c_float * reps)()
        rc = lib.calib_probe_run(op, blocks, iters, reps, ms)
        if rc != 0:
            raise SystemExit(f"calib_probe_run({op}) failed: HIP error {rc}")
        t = float(np.median(np.array(list(ms), dtype=np.float64))) * 1e-3
        rec = {"probe": PROBES[op], "waves_per_simd": waves_per_simd, "blocks": blocks, "median_ms": 1e3 * t}
        if op < 4:   # op 4's instruction count comes from the SQ_INSTS_VALU counter
            per_lane = lib.calib_fma_per_lane(iters) * (2 if op == 3 else 1)   # instructions
            wave_instr = per_lane * blocks * 256 / 64
            rec.update({"valu_instr_per_s": wave_instr / t, "valu_instr_per_cu_per_ns": wave_instr / t / cus / 1e9})
        if op == 0:
            rec["tflops"] = 2 * per_lane * blocks * 256 / t / 1e12
        out["probes"].append(rec)
    print(json.dumps(out), flush=True)
    return 0


def bench_learned_vs_plain(args):
    t_start = time.perf_counter()
    tet = learned_vs_plain_mesh("tet40", args)
    hexr = learned_vs_plain_mesh("hex48", args) if args.with_hex else None
    lb = tet["learned_batched"]
    if "error" in lb or not lb.get("reached_plain_quality"):
        value = 0.0
        lb_s = None
    else:
        lb_s = lb["learned_ms_per_step"] * lb["learned_iterations"] / 1e3
        value = tet["_plain_updates"] / lb_s
    K = args.steps
    roof_achieved = tet["_learned_flop_launch"] / (tet["_learned_kernel_us"] * 1e-6) / 1e12
    roof = {"bound": "valu", "achieved": roof_achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": roof_achieved / PEAK_FP32_TFLOPS, "traffic": WORKLOADS["tet40"].get("learned_traffic"),
            "kernel": "k_search_tet_learned_def<8,true> (network, proposal windows and the deferred full "
                      "searches of one colour class)",
            "mean_launch_us": tet["_learned_kernel_us"], "flop_per_launch": tet["_learned_flop_launch"],
            "flop_basis": "FP32 operations of the canonical evaluation per (candidate, ring element): "
                          f"{canon_flop('tet', 4):.2f} in the windows, {canon_flop('tet', 8):.3f} in the full "
                          "searches; the network's BF16 MFMA work is reported as mfma_tflops",
            # the executed view from the kernel's counter pass (not measured by this run)
            "valu_issue_frac_profiled": VALU_ISSUE_FRAC["k_search_tet_learned_def"],
            "valu_issue_basis": VALU_ISSUE_BASIS}
    cpu = None if args.no_cpu else cpu_baseline_tet(tet["_xyz"], tet["_cells"], "tet40")
    for r in (tet, hexr):
        if r:
            for k in [k for k in r if k.startswith("_")]:
                r.pop(k)
    out = {
        "metric": METRIC, "value": value, "unit": "vertex updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": (1e3 * lb_s / K) if lb_s else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (search), bf16 (network)", "data": "synthetic",
        "config": {
            "workload": ("learned + batched DMO against plain GPU DMO at equal final min and mean mean-ratio "
                         "quality (-1e-3), 40^3 x 6 tet cube (68,921 vertices, 384,000 tets, jitter 0.35h, seed "
                         f"{args.seed}, boundary fixed): value = the plain run's {K} iterations of vertex updates "
                         "(59,319 inner searches each) / the wall time of the learned + batched run to the plain "
                         "run's final quality (a one-mesh batch, one iteration replayed as a hipGraph); "
                         "ms_per_step = that time / K; the 48^3 hex cube beside it"),
            "name": "learned_vs_plain",
            "speedup_vs_plain": {"tet40_learned_batched": lb.get("speedup_vs_single_mesh_plain"),
                                 "tet40_learned": tet["learned"]["speedup_vs_plain"],
                                 "hex48_learned_batched": (hexr["learned_batched"].get("speedup_vs_single_mesh_plain")
                                                           if hexr else None),
                                 "hex48_learned": hexr["learned"]["speedup_vs_plain"] if hexr else None},
            "tet40": tet, "hex48": hexr,
            "bench_wall_s": time.perf_counter() - t_start,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="learned_vs_plain",
                    help="learned_vs_plain (default): learned + batched vs plain at equal quality on tet40 and hex48; "
                         "the others: one plain (or --mode learned) line per workload")
    ap.add_argument("--with-hex", type=int, default=1, help="learned_vs_plain: also the 48^3 hex cube (0: skip)")
    ap.add_argument("--calibrate", action="store_true",
                    help="measurement calibration instead of a bench line: a 256 MiB device copy (HBM counter "
                         "scale) and a kernel with a known count of FP32 FMA instructions (VALU issue rate)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="CPU baseline iterations (0: whole iterations until about 10 s of CPU time)")
    ap.add_argument("--tune-iters", type=int, default=3,
                    help="iterations per autotuner trial (batch256); 0 skips the tuner and sets the configuration "
                         "it keeps on the 256-mesh mix")
    ap.add_argument("--mode", choices=["plain", "learned"], default="plain")
    ap.add_argument("--waves", type=int, default=1, help="wavefronts per vertex search (1, 2, 4)")
    ap.add_argument("--compare-learned", type=int, default=None,
                    help="attach the learned-vs-plain comparison to the plain line (default: on for tet40)")
    ap.add_argument("--pairs", type=int, default=2_000_000, help="training pairs (learned mode)")
    ap.add_argument("--epochs", type=int, default=20)
    # proposal window (levels of 4^3, fraction of the plain first-level window);
    # default per element type, chosen on the seed-7 meshes, not the benchmark's
    ap.add_argument("--learn-depth", type=int, default=None, help="default: 1 (tet), 1 (hex)")
    ap.add_argument("--learn-frac", type=float, default=None, help="default: 0.025 (tet), 2.0 (hex)")
    ap.add_argument("--learn-mode", type=int, choices=[1, 3], default=None,
                    help="proposal of the learned runs: 1 the network's window, 3 the network's window and, where it "
                         "finds nothing better, a quarter-size window around the vertex before the full search "
                         "(default: 3 tet, 1 hex)")
    ap.add_argument("--learn-fallback", choices=["deferred", "immediate"], default="deferred",
                    help="fallback policy of the measured learned run (the other one is reported beside it)")
    ap.add_argument("--learned-extras", type=int, default=1,
                    help="1: also run the other fallback policy and the specified 1/4 window (0: skip them)")
    ap.add_argument("--learned-batched", type=int, default=1,
                    help="learned mode / comparison: also run the mesh as a one-mesh batch, learned vs plain (0: skip)")
    ap.add_argument("--batched-graph", type=int, default=1,
                    help="learned-batched comparison: 1 replays a captured iteration, 0 launches it eagerly")
    ap.add_argument("--batch-graph", type=int, default=1,
                    help="batch256: 1 replays a captured iteration, 0 launches it eagerly (same results)")
    ap.add_argument("--batch-learned", type=int, default=1,
                    help="batch256: also run the batch with the learned proposal on its tet meshes (0: skip)")
    # the batch's tet meshes span 500-50,000 vertices: twice the pairs (from
    # ~44 training meshes instead of ~22) and 30 epochs lowered the fallback
    # rate from 13 % to 9 % (on the tet40 cube the same change did not help)
    ap.add_argument("--batch-pairs", type=int, default=4_000_000, help="training pairs of the batch256 network")
    ap.add_argument("--batch-epochs", type=int, default=30)
    ap.add_argument("--batch-fallback", choices=["deferred", "immediate"], default="deferred",
                    help="fallback policy of the batch256 learned run")
    ap.add_argument("--tune-log", default=os.environ.get("VSDMO_TUNE_LOG", "autotune_batch256.jsonl"))
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this DMO build runs on a single MI355X and has no multi-GPU "
              "path; run it with --gpus 1.", file=sys.stderr)
        return 2
    if args.calibrate:
        return calibrate()
    if args.workload == "learned_vs_plain":
        return bench_learned_vs_plain(args)

    if args.compare_learned is None:
        args.compare_learned = 0          # the learned_vs_plain line carries the comparison
    hexw = WORKLOADS[args.workload]["kind"] == "hex"
    if args.learn_depth is None:
        args.learn_depth = LEARN_WINDOW["hex" if hexw else "tet"][0]
    if args.learn_frac is None:
        args.learn_frac = LEARN_WINDOW["hex" if hexw else "tet"][1]
    if args.learn_mode is None:
        args.learn_mode = LEARN_MODE["hex" if hexw else "tet"]
    if args.workload == "batch256":
        return bench_batch(args)
    if args.mode == "learned":
        return bench_learned(args)

    from volumesurfacediscretemeshoptimization_amd import (DMOHexClass, DMOMeshHexFactory, DMOMeshQuadFactory,
                                                           DMOMeshTetFactory, DMOMeshTriFactory, DMOQuadClass,
                                                           DMOTetClass, DMOTriClass, DMOTriFlatClass)

    wl = WORKLOADS[args.workload]
    kind = wl["kind"]
    surface_mesh = bool(wl.get("surface"))
    fac = {"tet": DMOMeshTetFactory, "hex": DMOMeshHexFactory, "tri": DMOMeshTriFactory, "quad": DMOMeshQuadFactory}[kind]
    cls = {"tet": DMOTetClass, "hex": DMOHexClass, "tri": DMOTriClass if surface_mesh else DMOTriFlatClass,
           "quad": DMOQuadClass}[kind]
    xyz, cells, desc, kw = make_mesh(args.workload, args.seed)
    kw["waves_per_vertex"] = args.waves
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
    searched = st["searches_inner"] if kind in ("tet", "hex") else st["searches_surface"]
    updates = st["searches_inner"] + st["searches_surface"] + st["searches_feature"]
    value = updates / dt
    flops = searched * wl["flop"] + (st["searches_feature"] * wl["flop1d"] if surface_mesh else 0)
    kernel_s = st["kernel_ms"] / 1e3
    launches = st["kernel_launches"]
    achieved = flops / kernel_s / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": wl["traffic"],
            "kernel": wl["kernel"], "launches": launches,
            "mean_launch_us": 1e3 * st["kernel_ms"] / max(launches, 1),
            "flop_per_update": wl["flop"], "cbrt_per_update": wl["cbrt"],
            "flop_basis": ("canonical FP32 operations of a full search (candidates x ring elements x operations "
                           "per pair)" if kind in ("tet", "hex") else "FLOP of the reference formulas")}
    if kind in ("tet", "hex") and not surface_mesh:
        roof["valu_issue_frac_profiled"] = VALU_ISSUE_FRAC[f"k_search_{kind}"]
        roof["valu_issue_basis"] = VALU_ISSUE_BASIS
    cpu = None
    if not args.no_cpu:
        threads = host_cpu()["threads_all"]
        cpu = cpu_baseline(kind, xyz, cells, args.cpu_iters, threads, not kw["fixed_boundary"] and not surface_mesh,
                           args.workload, surface_mesh)
    desc += f"; {args.waves} wavefront(s) per vertex"
    learned = None
    if args.compare_learned and kind in ("tet", "hex"):
        learned = learned_comparison(args)
        learned.pop("roofline", None)
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
    if learned is not None:
        out["config"]["learned_vs_plain"] = learned
    print(json.dumps(out), flush=True)
    opt.close()
    mesh.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
