"""GPU tests of the batched sweep, its hipGraph replay, the waves-per-vertex
variants and the autotuner.

Bars: every mesh of a batch ends bit-exact where one optimiser per mesh ends
(the batch only changes the schedule); the work order inside a colour round
and the waves per vertex change nothing; graph replays equal eager launches
after every iteration; the autotuned configuration keeps every mesh's final
minimum quality within 1e-3 of the (8, 3) run.
"""
import json
import os

import numpy as np
import pytest

from volumesurfacediscretemeshoptimization_amd import (DMOBatch, DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                       DMOMeshTriFactory, DMOTetClass, DMOTriFlatClass)
from volumesurfacediscretemeshoptimization_amd import synth

pytestmark = pytest.mark.gpu

FAC = {"tri": DMOMeshTriFactory, "tet": DMOMeshTetFactory, "hex": DMOMeshHexFactory}
CLS = {"tri": DMOTriFlatClass, "tet": DMOTetClass, "hex": DMOHexClass}


def _inputs():
    """A small uneven mix: flat tri grids, tet cubes and one hex cube."""
    out = [("tri", *synth.tri_grid(9, seed=1, jitter=0.35)),
           ("tet", *synth.tet_cube(7, seed=2)),
           ("tri", *synth.tri_grid(23, seed=3, jitter=0.35)),
           ("hex", *synth.hex_cube(6, seed=4, jitter=0.2)),
           ("tet", *synth.tet_cube(4, seed=5)),
           ("tri", *synth.tri_grid(5, seed=6, jitter=0.35)),
           ("tet", *synth.tet_cube(11, seed=7))]
    return out


def _meshes(inputs):
    return [FAC[k].create(xyz, cells) for k, xyz, cells in inputs]


def test_batch_matches_single_meshes():
    inputs = _inputs()
    iters = 3
    ref = []
    for k, xyz, cells in inputs:
        m = FAC[k].create(xyz, cells)
        o = CLS[k](m, n_iter=iters, fixed_boundary=True)
        o.iterate(iters)
        ref.append((m.positions(original_order=False), o.getQuality()))
        o.close()
        m.close()
    meshes = _meshes(inputs)
    b = DMOBatch(meshes)
    st0 = b.stats()
    assert st0["meshes"] == len(inputs)
    b.iterate(iters)
    q = b.min_quality()
    for i, (p, qm) in enumerate(ref):
        assert np.array_equal(b.positions(i), p), i
        assert q[i] == qm, (i, q[i], qm)
    st = b.stats()
    assert st["iterations"] == iters and st["searches"] == iters * st["items_per_iteration"]
    # write-back puts the batch positions into the meshes
    b.writeback()
    for i, (p, _) in enumerate(ref):
        assert np.array_equal(meshes[i].positions(original_order=False), p)
    b.close()


@pytest.mark.parametrize("kw", [{"shuffle_seed": 12345}, {"waves_per_vertex": 2}, {"waves_per_vertex": 4},
                                {"waves_per_vertex": 2, "shuffle_seed": 99, "use_graph": False},
                                {"persistent": False}, {"persistent": False, "waves_per_vertex": 2}])
def test_batch_schedule_independence(kw):
    inputs = _inputs()
    base = DMOBatch(_meshes(inputs))
    base.iterate(3)
    other = DMOBatch(_meshes(inputs), **kw)
    other.iterate(3)
    for i in range(len(inputs)):
        assert np.array_equal(base.positions(i), other.positions(i)), (kw, i)
    assert np.array_equal(base.min_quality(), other.min_quality())
    base.close()
    other.close()


def test_batch_graph_replay_equals_eager():
    """Graph replays interleaved with eager launches and host syncs: every
    replay's result equals the eager one."""
    inputs = _inputs()
    eager = DMOBatch(_meshes(inputs), use_graph=False)
    graph = DMOBatch(_meshes(inputs), use_graph=True)
    for it in range(4):
        eager.iterate(1)
        graph.iterate(1)
        for i in range(len(inputs)):
            assert np.array_equal(eager.positions(i), graph.positions(i)), (it, i)
        assert np.array_equal(eager.min_quality(), graph.min_quality()), it
    # several replays in one call
    eager.iterate(3)
    graph.iterate(3)
    for i in range(len(inputs)):
        assert np.array_equal(eager.positions(i), graph.positions(i)), i
    eager.close()
    graph.close()


def test_batch_restore_and_params():
    inputs = _inputs()
    b = DMOBatch(_meshes(inputs))
    q0 = b.min_quality()
    p0 = [b.positions(i) for i in range(len(inputs))]
    b.iterate(2)
    assert np.all(b.min_quality() >= q0)
    b.restore()
    for i in range(len(inputs)):
        assert np.array_equal(b.positions(i), p0[i])
    assert np.array_equal(b.min_quality(), q0)
    b.set_params(waves_per_vertex=2)
    assert b.params()["waves_per_vertex"] == 2
    with pytest.raises(Exception):
        b.set_params(nq=6, waves_per_vertex=4)
    b.close()


def test_autotune_keeps_quality(tmp_path):
    inputs = _inputs()
    iters = 4
    tol = 1e-3
    ref = DMOBatch(_meshes(inputs))
    ref.iterate(iters)
    qref = ref.min_quality()
    ref.close()
    b = DMOBatch(_meshes(inputs))
    log = tmp_path / "tune.jsonl"
    chosen = b.autotune(iters=iters, tol=tol, log_path=str(log))
    lines = [json.loads(x) for x in log.read_text().splitlines()]
    assert lines[-1]["stage"] == "chosen"
    assert lines[-1]["nq"] == chosen["nq"] and lines[-1]["depth"] == chosen["depth"]
    assert sum(1 for x in lines if x["stage"] == "schedule") == 6
    assert b.stats()["iterations"] == 0          # the tuner leaves no trace
    b.iterate(iters)
    q = b.min_quality()
    assert np.all(q >= qref - tol), (chosen, q - qref)
    b.close()


@pytest.mark.parametrize("kind,w", [("tet", 2), ("tet", 4), ("hex", 2), ("hex", 4), ("tri", 2)])
def test_single_mesh_waves_per_vertex(kind, w):
    gen = {"tet": lambda: synth.tet_cube(10, seed=8), "hex": lambda: synth.hex_cube(7, seed=8),
           "tri": lambda: synth.tri_grid(20, seed=8)}[kind]
    xyz, cells = gen()
    m1 = FAC[kind].create(xyz, cells)
    m2 = FAC[kind].create(xyz, cells)
    o1 = CLS[kind](m1, n_iter=3, fixed_boundary=True)
    o2 = CLS[kind](m2, n_iter=3, waves_per_vertex=w, fixed_boundary=True)
    o1.iterate(3)
    o2.iterate(3)
    assert np.array_equal(m1.positions(original_order=False), m2.positions(original_order=False))
    o1.close()
    o2.close()


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_split_launch_bit_exact(kind):
    """waves_per_vertex = 0: one vertex per wave with a 4-wave tail per class
    (classes here have ~1,500 vertices: head and tail both used)."""
    gen = {"tet": lambda: synth.tet_cube(24, seed=9), "hex": lambda: synth.hex_cube(24, seed=9, jitter=0.2)}[kind]
    xyz, cells = gen()
    m1 = FAC[kind].create(xyz, cells)
    m2 = FAC[kind].create(xyz, cells)
    off = m1.color_offsets("inner")
    assert np.diff(off).max() > 1024
    o1 = CLS[kind](m1, n_iter=2, fixed_boundary=True, waves_per_vertex=1)
    o2 = CLS[kind](m2, n_iter=2, fixed_boundary=True, waves_per_vertex=0)
    o1.iterate(2)
    o2.iterate(2)
    assert np.array_equal(m1.positions(original_order=False), m2.positions(original_order=False))
    o1.close()
    o2.close()
