"""GPU parity of the learned proposal against the oracle's restatement of it.

The oracle runs the learned iteration class by class on the CPU: each
class's one-ring features (tests/mlp_ref.py, bit-exact with the GPU's), the
network's outputs for them (vsdmo_mlp_forward, the network the search
kernels run inside them, bit for bit), the proposal c = v + L y, the window
search (or_learned_window, pinned against tests/ref64.py; mode 3 searches a
window a quarter that size around v where the network's finds nothing better),
and the plain full
search (or_dmo_vertex) for the vertices that fall back -- in the same class
pass (immediate policy) or in the class's next pass (deferred).  The GPU
optimiser and a one-mesh batch must end every iteration on the same
positions, bit for bit, with the same fallback count.
"""
import numpy as np
import pytest

import mlp_ref
from volumesurfacediscretemeshoptimization_amd import (DMOBatch, DMODataset, DMOHexClass, DMOMeshHexFactory,
                                                       DMOMeshTetFactory, DMOMLP, DMOTetClass, synth)
from volumesurfacediscretemeshoptimization_amd import _lib

pytestmark = pytest.mark.gpu

ITERS = 3


def _mesh(kind, n, seed):
    return synth.tet_cube(n, seed=seed) if kind == "tet" else synth.hex_cube(n, seed=seed, jitter=0.2)


@pytest.fixture(scope="module")
def nets():
    """a briefly trained network per element type (pairs from other seeds)"""
    out = {}
    for kind, fac, cls, n in (("tet", DMOMeshTetFactory, DMOTetClass, 10), ("hex", DMOMeshHexFactory, DMOHexClass, 8)):
        ds = DMODataset(1 << 18)
        for seed in (300, 301):
            xyz, cells = _mesh(kind, n, seed)
            m = fac.create(xyz, cells)
            o = cls(m, n_iter=3, fixed_boundary=True)
            o.record(ds, 3)
            o.close()
            m.close()
        mlp = DMOMLP(seed=5)
        mlp.train(ds, epochs=4, batch=512, lr=1e-3, seed=6)
        ds.close()
        out[kind] = mlp
    yield out
    for m in out.values():
        m.close()


def oracle_learned(oracle, kind, xyz, cells, mlp, mode, depth_l, frac, policy, iters, hex_metric=0):
    """The learned iterations restated on the CPU; returns the positions after
    each iteration and the number of full searches run for fallbacks (under
    the deferred policy a deferral counts when its search runs, in the next
    iteration, as the optimiser counts it)."""
    om = oracle.OracleMesh(kind, xyz, cells)
    om.hex_metric = hex_metric
    row, val = om.ring()
    stride = 3 if kind == "tet" else 7
    off = om.offsets("inner")
    nbr = {v: mlp_ref.canonical_neighbours(row, val, stride, v) for v in range(int(off[0]), int(off[-1]))}
    epoch = {}
    fallbacks = 0
    hist = []
    for it in range(iters):
        for c in range(len(off) - 1):
            lo, hi = int(off[c]), int(off[c + 1])
            pos = om.positions()
            new = pos.copy()
            if mode in (1, 3):
                feats, ls = zip(*(mlp_ref.features(pos, nbr[v], v) for v in range(lo, hi)))
                Y = mlp.forward(np.stack(feats)).astype(np.float32)
            for i, v in enumerate(range(lo, hi)):
                if policy == "deferred" and epoch.get(v) == it - 1:
                    _, out = om.search_vertex(v, pos)          # the deferred full search
                    new[v] = out[v]
                    fallbacks += 1
                    continue
                x0 = pos[v]
                c3 = (x0 + np.float32(ls[i]) * Y[i]).astype(np.float32) if mode in (1, 3) else x0
                st, p = om.learned_window(v, pos, c3, depth_l, frac)
                if mode == 3 and st == -1:                 # a quarter-size window around the vertex itself
                    st, p = om.learned_window(v, pos, x0, depth_l, float(np.float32(frac) * np.float32(0.25)))
                if st == 1:
                    new[v] = p
                elif st == -1:
                    if policy == "immediate":
                        _, out = om.search_vertex(v, pos)
                        new[v] = out[v]
                        fallbacks += 1
                    else:
                        epoch[v] = it
            om.set_positions(new)
        hist.append(om.positions().copy())
    return hist, fallbacks


CASES = [
    # kind, mode, levels, fraction of the plain first-level window, policy
    ("tet", 1, 1, 0.025, "deferred"),      # the shipped tet window, bench policy
    ("tet", 1, 1, 0.025, "immediate"),
    ("tet", 1, 1, 0.25, "deferred"),       # the specified 1/4 window
    ("tet", 2, 2, 0.5, "immediate"),       # no network, two levels
    ("hex", 1, 1, 2.0, "deferred"),        # the shipped hex window
    ("hex", 1, 1, 2.0, "immediate"),
    ("hex", 2, 1, 0.25, "deferred"),
    ("tet", 3, 1, 0.025, "deferred"),      # network, then the vertex's own window
    ("tet", 3, 2, 0.05, "immediate"),
    ("hex", 3, 1, 2.0, "deferred"),
]


@pytest.mark.parametrize("kind,mode,depth,frac,policy", CASES)
def test_learned_iterations_match_oracle(oracle, nets, kind, mode, depth, frac, policy):
    n = 10 if kind == "tet" else 8
    xyz, cells = _mesh(kind, n, 17)
    mlp = nets[kind] if mode in (1, 3) else None
    ref, ref_fb = oracle_learned(oracle, kind, xyz, cells, mlp, mode, depth, frac, policy, ITERS)
    fac, cls = (DMOMeshTetFactory, DMOTetClass) if kind == "tet" else (DMOMeshHexFactory, DMOHexClass)
    m = fac.create(xyz, cells)
    o = cls(m, n_iter=ITERS, fixed_boundary=True)
    try:
        o.set_proposal(mlp, mode=mode, depth=depth, window_frac=frac, fallback=policy)
        for it in range(ITERS):
            o.iterate(1)
            gp = m.positions(original_order=False)
            bad = np.flatnonzero((gp != ref[it]).any(axis=1))
            assert bad.size == 0, f"iteration {it}: {bad.size} vertices differ, first {bad[:5]}"
        st = o.stats()
        assert st["fallbacks"] == ref_fb, (st["fallbacks"], ref_fb)
        assert ref_fb < st["searches_learned"]
        if mode == 1:
            assert ref_fb > 0            # the trained network's proposals fall back now and then
    finally:
        o.close()
        m.close()
    # the same mesh as a one-mesh batch
    mb = fac.create(xyz, cells)
    b = DMOBatch([mb], use_graph=True)
    try:
        b.set_fallback(policy)
        b.set_proposal(mlp, mode=mode, depth=depth, window_frac=frac,
                       elem=_lib.VSDMO_TET if kind == "tet" else _lib.VSDMO_HEX)
        for it in range(ITERS):
            b.iterate(1)
            assert np.array_equal(b.positions(0), ref[it]), f"batch, iteration {it}"
        assert b.learned_stats()["fallbacks"] == ref_fb
    finally:
        b.close()
        mb.close()


def test_learned_hex_corner8_matches_oracle(oracle, nets):
    """the learned hex iteration with the eight-corner metric (windows,
    fallbacks and the deferred full searches all score on it), single mesh
    and one-mesh batch"""
    xyz, cells = _mesh("hex", 8, 18)
    ref, ref_fb = oracle_learned(oracle, "hex", xyz, cells, nets["hex"], 1, 1, 2.0, "deferred", ITERS,
                                 hex_metric=oracle.OR_HEX_CORNER8)
    m = DMOMeshHexFactory.create(xyz, cells)
    o = DMOHexClass(m, n_iter=ITERS, fixed_boundary=True, hex_metric=_lib.VSDMO_HEX_CORNER8)
    try:
        o.set_proposal(nets["hex"], mode=1, depth=1, window_frac=2.0, fallback="deferred")
        for it in range(ITERS):
            o.iterate(1)
            assert np.array_equal(m.positions(original_order=False), ref[it]), it
        assert o.stats()["fallbacks"] == ref_fb
    finally:
        o.close()
        m.close()
    mb = DMOMeshHexFactory.create(xyz, cells)
    b = DMOBatch([mb], use_graph=True, hex_metric=_lib.VSDMO_HEX_CORNER8)
    try:
        b.set_fallback("deferred")
        b.set_proposal(nets["hex"], mode=1, depth=1, window_frac=2.0, elem=_lib.VSDMO_HEX)
        for it in range(ITERS):
            b.iterate(1)
            assert np.array_equal(b.positions(0), ref[it]), f"batch, iteration {it}"
    finally:
        b.close()
        mb.close()


def _icosphere():
    """42 unit vectors and 80 triangles: the icosahedron subdivided once"""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    pts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    mid = {}

    def m(a, b):
        k = (min(a, b), max(a, b))
        if k not in mid:
            p = pts[a] + pts[b]
            pts.append(p / np.linalg.norm(p))
            mid[k] = len(pts) - 1
        return mid[k]
    tri = []
    for a, b, c in f:
        ab, bc, ca = m(a, b), m(b, c), m(c, a)
        tri += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
    return np.array(pts), np.array(tri)


def star_mesh(seed=3):
    """A vertex with 42 distinct one-ring neighbours (more than the network's
    32-entry neighbour table holds): the centre of two icosphere shells (r =
    0.5, 1; the inner shell's vertices are interior too), tets from the centre
    to the inner shell and prisms between the shells split into 3 tets by
    vertex index; positions jittered, tets oriented positive."""
    s, tri = _icosphere()
    n = len(s)
    rng = np.random.default_rng(seed)
    xyz = np.concatenate([np.zeros((1, 3)), 0.5 * s, s])
    xyz[: 1 + n] += rng.uniform(-0.04, 0.04, size=(1 + n, 3))
    cells = []
    for a, b, c in tri:
        cells.append((0, 1 + a, 1 + b, 1 + c))
        i, j, k = sorted((a, b, c))
        lo = lambda q: 1 + q
        hi = lambda q: 1 + n + q
        cells += [(lo(i), lo(j), lo(k), hi(k)), (lo(i), lo(j), hi(j), hi(k)), (lo(i), hi(i), hi(j), hi(k))]
    cells = np.array(cells, np.int32)
    p = xyz[cells]
    vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    cells[vol < 0, 1], cells[vol < 0, 2] = cells[vol < 0, 2].copy(), cells[vol < 0, 1].copy()
    return xyz.astype(np.float32), cells


@pytest.mark.parametrize("mode", [1, 3])
def test_learned_truncated_neighbour_table_matches_oracle(oracle, nets, mode):
    """The window kernel stages a ring corner from the network's neighbour
    table when the table holds it, and gathers it otherwise: the star mesh's
    centre has 42 distinct neighbours, 10 of which are outside its table."""
    xyz, cells = star_mesh()
    centre_nbrs = set(cells[(cells == 0).any(axis=1)].ravel().tolist()) - {0}
    assert len(centre_nbrs) > 32
    ref, ref_fb = oracle_learned(oracle, "tet", xyz, cells, nets["tet"], mode, 1, 0.25, "deferred", ITERS)
    m = DMOMeshTetFactory.create(xyz, cells)
    o = DMOTetClass(m, n_iter=ITERS, fixed_boundary=True)
    try:
        p0 = m.positions(original_order=False)
        o.set_proposal(nets["tet"], mode=mode, depth=1, window_frac=0.25, fallback="deferred")
        for it in range(ITERS):
            o.iterate(1)
            gp = m.positions(original_order=False)
            bad = np.flatnonzero((gp != ref[it]).any(axis=1))
            assert bad.size == 0, f"iteration {it}: {bad.size} vertices differ, first {bad[:5]}"
        assert o.stats()["fallbacks"] == ref_fb
        assert (m.positions(original_order=False) != p0).any(axis=1).sum() > 0   # the search moved vertices
    finally:
        o.close()
        m.close()
    # the same mesh as a one-mesh batch (its ring-corner table positions copied into the batch)
    mb = DMOMeshTetFactory.create(xyz, cells)
    b = DMOBatch([mb], use_graph=True)
    try:
        b.set_fallback("deferred")
        b.set_proposal(nets["tet"], mode=mode, depth=1, window_frac=0.25, elem=_lib.VSDMO_TET)
        for it in range(ITERS):
            b.iterate(1)
            assert np.array_equal(b.positions(0), ref[it]), f"batch, iteration {it}"
        assert b.learned_stats()["fallbacks"] == ref_fb
    finally:
        b.close()
        mb.close()
