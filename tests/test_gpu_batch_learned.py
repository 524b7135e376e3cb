"""GPU tests of the learned proposal inside the batched sweep.

Bars: every tet (hex) mesh of a batch with a tet (hex) proposal ends
bit-exact where one DMOTetClass (DMOHexClass) per mesh (boundary fixed)
with the same network, window and the immediate fallback policy ends, with
the same fallback count; the other meshes of the same batch end where their
plain optimisers end; graph replays and shuffled work orders equal eager
launches in mesh order; switching the proposal off gives plain searches
again.
"""
import numpy as np
import pytest

from volumesurfacediscretemeshoptimization_amd import (DMOBatch, DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                       DMOMeshTriFactory, DMOMLP, DMOTetClass, DMOTriFlatClass,
                                                       VsdmoError)
from volumesurfacediscretemeshoptimization_amd import _lib, synth

pytestmark = pytest.mark.gpu

FAC = {"tri": DMOMeshTriFactory, "tet": DMOMeshTetFactory, "hex": DMOMeshHexFactory}
CLS = {"tri": DMOTriFlatClass, "tet": DMOTetClass, "hex": DMOHexClass}


def _inputs():
    """Tet cubes of uneven sizes between flat tri grids and one hex cube, and
    a one-cell tet cube whose 8 vertices are all on the boundary (no items)."""
    return [("tet", *synth.tet_cube(9, seed=41)),
            ("tri", *synth.tri_grid(11, seed=42, jitter=0.35)),
            ("tet", *synth.tet_cube(5, seed=43)),
            ("hex", *synth.hex_cube(6, seed=44, jitter=0.2)),
            ("tet", *synth.tet_cube(1, seed=48)),
            ("tet", *synth.tet_cube(12, seed=45)),
            ("tri", *synth.tri_grid(6, seed=46, jitter=0.35))]


def _meshes(inputs):
    return [FAC[k].create(xyz, cells) for k, xyz, cells in inputs]


ELEM = {"tet": _lib.VSDMO_TET, "hex": _lib.VSDMO_HEX}


def _network(seed=21, b3=(0.05, -0.03, 0.02)):
    mlp = DMOMLP(seed=seed)
    p = mlp.params()
    p[-3:] = b3                          # a nonzero mean displacement
    mlp.set_params(p)
    return mlp


@pytest.mark.parametrize("types,mode,depth,frac", [(("tet",), 1, 1, 0.25), (("tet",), 1, 2, 0.025),
                                                   (("tet",), 2, 2, 0.25), (("hex",), 1, 1, 2.0),
                                                   (("tet", "hex"), 1, 1, 0.25), (("tet", "hex"), 2, 1, 0.5)])
def test_batch_learned_matches_single_meshes(types, mode, depth, frac):
    nets = {k: (_network(seed=21 + i, b3=(0.05, -0.03, 0.02 + 0.01 * i)) if mode == 1 else None)
            for i, k in enumerate(types)}
    inputs = _inputs()
    iters = 3
    ref, falls, visits = [], 0, 0
    for k, xyz, cells in inputs:
        m = FAC[k].create(xyz, cells)
        o = CLS[k](m, n_iter=iters, fixed_boundary=True)
        if k in types:
            o.set_proposal(nets[k], mode=mode, depth=depth, window_frac=frac, fallback="immediate")
            visits += iters * int(np.diff(m.color_offsets("inner")).sum())
        o.iterate(iters)
        if k in types:
            falls += o.stats()["fallbacks"]
        ref.append((m.positions(original_order=False), o.getQuality()))
        o.close()
        m.close()
    meshes = _meshes(inputs)
    b = DMOBatch(meshes)
    for k in types:
        b.set_proposal(nets[k], mode=mode, depth=depth, window_frac=frac, elem=ELEM[k])
    b.iterate(iters)
    q = b.min_quality()
    for i, (p, qm) in enumerate(ref):
        assert np.array_equal(b.positions(i), p), (inputs[i][0], i)
        assert q[i] == qm, (i, q[i], qm)
    ls = b.learned_stats()
    assert ls["learned_visits"] == visits
    assert ls["fallbacks"] == falls and falls < visits, (ls, falls)
    b.close()
    for n in nets.values():
        if n is not None:
            n.close()


def test_batch_learned_graph_and_order_independence():
    """Graph replays (interleaved with host reads) and a shuffled work order
    give the eager, mesh-order positions after every iteration."""
    mlp, mlp_h = _network(), _network(seed=23)
    inputs = _inputs()
    eager = DMOBatch(_meshes(inputs), use_graph=False)
    graph = DMOBatch(_meshes(inputs), use_graph=True, shuffle_seed=777, waves_per_vertex=2)
    for b in (eager, graph):
        b.set_proposal(mlp, mode=1)
        b.set_proposal(mlp_h, mode=1, elem=_lib.VSDMO_HEX)
    for it in range(3):
        eager.iterate(1)
        graph.iterate(1)
        for i in range(len(inputs)):
            assert np.array_equal(eager.positions(i), graph.positions(i)), (it, i)
        assert np.array_equal(eager.min_quality(), graph.min_quality()), it
    eager.iterate(2)
    graph.iterate(2)
    for i in range(len(inputs)):
        assert np.array_equal(eager.positions(i), graph.positions(i)), i
    assert eager.learned_stats() == graph.learned_stats()
    eager.close()
    graph.close()
    mlp.close()
    mlp_h.close()


def test_batch_proposal_off_and_refusals():
    """Mode 0 after a learned setting gives the plain batch's positions; the
    autotuner refuses while a proposal is set; bad arguments are refused."""
    mlp = _network()
    inputs = _inputs()
    plain = DMOBatch(_meshes(inputs))
    plain.iterate(2)
    b = DMOBatch(_meshes(inputs))
    b.set_proposal(mlp, mode=1)
    with pytest.raises(VsdmoError):
        b.autotune(iters=1)
    with pytest.raises(VsdmoError):
        b.set_proposal(None, mode=1)
    with pytest.raises(VsdmoError):
        b.set_proposal(mlp, mode=1, depth=0)
    with pytest.raises(VsdmoError):
        b.set_proposal(mlp, mode=1, elem=_lib.VSDMO_TRI)
    b.set_proposal(mlp, mode=2, elem=_lib.VSDMO_HEX)
    with pytest.raises(VsdmoError):
        b.autotune(iters=1)
    b.set_proposal(None, mode=0, elem=_lib.VSDMO_HEX)
    b.set_proposal(None, mode=0)
    b.iterate(2)
    for i in range(len(inputs)):
        assert np.array_equal(plain.positions(i), b.positions(i)), i
    assert b.learned_stats()["learned_visits"] == 0
    plain.close()
    b.close()
    mlp.close()


def test_network_outlives_its_users():
    """vsdmo_mlp_destroy refuses while an optimiser or batch holds the network
    as its proposal, and succeeds once the proposal is off or the holder is
    destroyed."""
    mlp = _network()
    xyz, cells = synth.tet_cube(5, seed=47)
    m = DMOMeshTetFactory.create(xyz, cells)
    o = DMOTetClass(m, n_iter=1, fixed_boundary=True)
    o.set_proposal(mlp, mode=1)
    b = DMOBatch([DMOMeshTetFactory.create(xyz, cells)])
    b.set_proposal(mlp, mode=1)
    with pytest.raises(VsdmoError):
        mlp.close()
    o.set_proposal(None, mode=0)
    with pytest.raises(VsdmoError):
        mlp.close()               # still held by the batch
    b.close()
    mlp.close()
    # a proposal switched off leaves a working plain optimiser
    o.iterate(1)
    o.close()
    m.close()


@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("types", [("tet",), ("hex",), ("tet", "hex")])
def test_batch_deferred_matches_single_meshes(types, persistent):
    """Deferred fallback in the batch (iteration number and double-buffered
    lists on the device, one captured iteration replayed): every learned mesh
    ends bit-exact where a single-mesh optimiser with the deferred policy
    ends, with the same fallback count, after every iteration; the plain
    items run on the persistent queue or one block per item (one launch per
    element type)."""
    nets = {k: _network(seed=31 + i, b3=(0.04, 0.03, -0.02 - 0.01 * i)) for i, k in enumerate(types)}
    inputs = _inputs()
    iters = 4
    ref, falls = [], 0
    for k, xyz, cells in inputs:
        m = FAC[k].create(xyz, cells)
        o = CLS[k](m, n_iter=iters, fixed_boundary=True)
        if k in types:
            o.set_proposal(nets[k], mode=1, depth=1, window_frac=0.25, fallback="deferred")
        rows = []
        for _ in range(iters):
            o.doIteration()
            rows.append(m.positions(original_order=False))
        if k in types:
            falls += o.stats()["fallbacks"]
        ref.append(rows)
        o.close()
        m.close()
    b = DMOBatch(_meshes(inputs), persistent=persistent)
    b.set_fallback("deferred")
    for k in types:
        b.set_proposal(nets[k], mode=1, depth=1, window_frac=0.25, elem=ELEM[k])
    for it in range(iters):
        b.iterate(1)
        for i, rows in enumerate(ref):
            assert np.array_equal(b.positions(i), rows[it]), (inputs[i][0], i, it)
    ls = b.learned_stats()
    assert ls["fallbacks"] == falls and falls > 0, (ls, falls)
    # restore clears the pending deferrals: the same iterations again
    b.restore()
    b.iterate(iters)
    for i, rows in enumerate(ref):
        assert np.array_equal(b.positions(i), rows[-1]), (inputs[i][0], i)
    b.close()
    for n in nets.values():
        n.close()


def test_batch_deferred_graph_equals_eager_and_refusals():
    mlp, mlp_h = _network(), _network(seed=23)
    inputs = _inputs()
    eager = DMOBatch(_meshes(inputs), use_graph=False)
    graph = DMOBatch(_meshes(inputs), use_graph=True, shuffle_seed=99)
    for b in (eager, graph):
        b.set_fallback("deferred")
        b.set_proposal(mlp, mode=1)
        b.set_proposal(mlp_h, mode=1, elem=_lib.VSDMO_HEX)
    for it in range(3):
        eager.iterate(1)
        graph.iterate(2 if it == 1 else 1)
        if it == 1:
            eager.iterate(1)
        for i in range(len(inputs)):
            assert np.array_equal(eager.positions(i), graph.positions(i)), (it, i)
    assert eager.learned_stats() == graph.learned_stats()
    with pytest.raises(VsdmoError):
        graph.set_params(nq=6)           # the deferred full searches need nq 4 or 8
    with pytest.raises(ValueError):
        graph.set_fallback("later")
    graph.set_fallback("immediate")
    graph.set_params(nq=6)
    for b in (eager, graph):
        b.close()
    mlp.close()
    mlp_h.close()
