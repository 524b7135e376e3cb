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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "vertex updates/s at equal final min mean-ratio quality; learned vs plain DMO"
CPU_BUDGET_S = 10.0
# learned proposal window per element type: (levels, fraction of the plain
# first-level window), chosen by sweeps with the deferred fallback on the
# seed-7 and seed-8 40^3 tet / 48^3 hex cubes (the benchmark meshes use seed 0)
LEARN_WINDOW = {"tet": (1, 0.025), "hex": (1, 2.0)}
# the learned configuration as specified: one 4x4x4 grid on 1/4 of the plain
# first-level window, reported beside the tuned window
SPEC_WINDOW = (1, 0.25)
PEAK_FP32_TFLOPS = 157.3
# Algorithmic work of one vertex update: (nq^d + 1) ring evaluations per level
# x levels + the final check = 1,540 (3D) / 196 (2D), times the one-ring size
# (24 tets / 8 hexes / 6 triangles) times the per-element FLOP of the metric
# (tet 73, hex 407, tri 23), cube roots counted separately.
# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE x 1.99981 copy-kernel
# calibration + WRITE_SIZE x 1.0, mean over the dispatches of 3 iterations,
# profiles/k_search_tet_tet40.md, profiles/k_search_hex_hex48.md,
# profiles/k_search_fan2d_surf.md; the learned pass over the headline run,
# profiles/learned_tet40.md and profiles/learned_hex48.md).
# Surface-mesh searches (algorithmic FLOP per vertex update, valence n = 6
# triangles / 4 quads): per candidate the projection onto the blended local
# surfaces (57 + 69 n) plus the signed fan metric (59 n triangles, 104 n
# quads); 3 levels x 64 candidates (2D) or 3 x 8 (1D, projection 45 + 52 * 2
# curve neighbours).
SURF_FLOP = {"tri": 3 * 64 * (57 + 69 * 6 + 59 * 6), "quad": 3 * 64 * (57 + 69 * 4 + 104 * 4)}
SURF_FLOP1D = {"tri": 3 * 8 * (149 + 59 * 6), "quad": 3 * 8 * (149 + 104 * 4)}
WORKLOADS = {
    "tet40": {"kind": "tet", "flop": 2_698_080, "cbrt": 36_960, "kernel": "k_search_tet<8>",
              "traffic": 10_970_146, "learned_traffic": 8_042_968},
    "hex48": {"kind": "hex", "flop": 5_014_240, "cbrt": 110_880, "kernel": "k_search_hex<8>",
              "traffic": 17_570_841, "learned_traffic": 12_274_568},
    # the 48^3 hex cube at jitter 0.15h: a valid mesh, so the surface and
    # feature passes run every iteration (the roofline is the inner kernel's)
    "hex48_surface": {"kind": "hex", "flop": 5_014_240, "cbrt": 110_880, "kernel": "k_search_hex<8>",
                      "traffic": None},
    "tri64": {"kind": "tri", "flop": 27_048, "cbrt": 0, "kernel": "k_search_flat<8>", "traffic": None},
    "surf_tri": {"kind": "tri", "surface": True, "flop": SURF_FLOP["tri"], "flop1d": SURF_FLOP1D["tri"], "cbrt": 0,
                 "kernel": "k_search_fan2d<false>", "traffic": 18_791_143},
    "surf_quad": {"kind": "quad", "surface": True, "flop": SURF_FLOP["quad"], "flop1d": SURF_FLOP1D["quad"], "cbrt": 0,
                  "kernel": "k_search_fan2d<true>", "traffic": 20_562_741},
    "batch256": {"kind": "batch", "flop": 2_698_080, "cbrt": 36_960, "kernel": "k_batch_round", "traffic": None},
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
    b.set_proposal(mlp, mode=1, depth=depth, window_frac=frac)
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
    chosen = b.autotune(iters=args.tune_iters, tol=1e-3, log_path=args.tune_log)
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
    # algorithmic work per iteration: tet searches x tet FLOP + tri searches x tri FLOP
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
        threads = min(16, os.cpu_count() or 1)
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
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "k_batch_round (all rounds + "
                     "per-mesh quality, device time of the replayed iterations)"},
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
        threads = min(16, os.cpu_count() or 1)
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
    b.set_proposal(mlp, mode=1, depth=args.learn_depth, window_frac=args.learn_frac, elem=elem)
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
    return {"policy": args.learn_fallback, "plain_iterations": args.steps, "learned_iterations": n,
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
        wo.set_proposal(mlp, mode=1, depth=args.learn_depth, window_frac=args.learn_frac, fallback=pol)
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
        ol.set_proposal(mlp, mode=1, depth=depth, window_frac=frac, fallback=policy)
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
    ok.set_proposal(mlp, mode=1, depth=args.learn_depth, window_frac=args.learn_frac, fallback=args.learn_fallback)
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
    ring = 24 if kind == "tet" else 8
    per_elem = 73 if kind == "tet" else 407
    evals = stl["searches_learned"] * (64 * args.learn_depth + 1) + stl["fallbacks"] * 1540
    flops = evals * ring * per_elem
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="tet40")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="CPU baseline iterations (0: whole iterations until about 10 s of CPU time)")
    ap.add_argument("--tune-iters", type=int, default=3, help="iterations per autotuner trial (batch256)")
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

    if args.compare_learned is None:
        args.compare_learned = int(args.workload == "tet40" and args.mode == "plain")
    hexw = WORKLOADS[args.workload]["kind"] == "hex"
    if args.learn_depth is None:
        args.learn_depth = LEARN_WINDOW["hex" if hexw else "tet"][0]
    if args.learn_frac is None:
        args.learn_frac = LEARN_WINDOW["hex" if hexw else "tet"][1]
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
            "flop_per_update": wl["flop"], "cbrt_per_update": wl["cbrt"]}
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
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
