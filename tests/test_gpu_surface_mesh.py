"""GPU parity of the surface-mesh path (triangle / quad surfaces through
DMOTriClass / DMOQuadClass, reference dmo/DMOTri.cu, dmo/DMOQuad.cu) and of
Laplace smoothing (dmo/Laplace.cu) against the oracle, through the C ABI.

Bars: mesh structures bit-exact; element qualities and positions after every
iteration bit-exact (the fan metric, the projections, the nearest-neighbour
walk and the normal updates are literal FP32 with contraction off on both
sides).
"""
import numpy as np
import pytest

from volumesurfacediscretemeshoptimization_amd import (DMOBatch, DMOMeshQuadFactory, DMOMeshTriFactory, DMOQuadClass,
                                                       DMOTriClass, DMOTriFlatClass, VsdmoError, laplaceSmoothing)
from volumesurfacediscretemeshoptimization_amd import synth

pytestmark = pytest.mark.gpu

CASES = {
    "tri_cube": lambda: ("tri",) + synth.cube_surface(10, seed=3, quads=False),
    "quad_cube": lambda: ("quad",) + synth.cube_surface(10, seed=3),
    "tri_field": lambda: ("tri",) + synth.height_field(16, seed=4, quads=False),
    "quad_field": lambda: ("quad",) + synth.height_field(16, seed=4),
    "tri_sphere": lambda: ("tri",) + synth.cube_surface(12, seed=5, quads=False, sphere=True),
    "quad_sphere": lambda: ("quad",) + synth.cube_surface(12, seed=5, sphere=True),
}


def _mesh(kind, xyz, f):
    return (DMOMeshQuadFactory if kind == "quad" else DMOMeshTriFactory).create(xyz, f)


def _opt(kind, g, **kw):
    return (DMOQuadClass if kind == "quad" else DMOTriClass)(g, **kw)


@pytest.mark.parametrize("case", sorted(CASES))
def test_surface_mesh_build_bitexact(oracle, case):
    kind, xyz, f = CASES[case]()
    g = _mesh(kind, xyz, f)
    o = oracle.OracleMesh(kind, xyz, f)
    c = g.counts()
    assert (c["n_surface"], c["n_free"], c["n_feature"], c["flat"]) == (o.m.nSurf, o.m.nFree, o.m.nFeature, o.m.flat)
    assert np.array_equal(g.permutation(), o.new_of_old())
    assert np.array_equal(g.cells(), o.cells())
    for which in ("free", "feature"):
        assert np.array_equal(g.color_offsets(which), o.offsets(which)), which
    grow, ghe = g.halfedges()
    orow, ohe = o.halfedges()
    assert np.array_equal(grow, orow) and np.array_equal(ghe, ohe)
    gb, gf = g.flags()
    ob, of = o.flags()
    assert np.array_equal(gb, ob) and np.array_equal(gf, of)
    g.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_surface_mesh_quality_bitexact(oracle, case):
    kind, xyz, f = CASES[case]()
    g = _mesh(kind, xyz, f)
    opt = _opt(kind, g, n_iter=1)
    o = oracle.OracleMesh(kind, xyz, f)
    assert np.array_equal(opt.getElementQualities().view(np.uint32), o.surface_mesh_qualities().view(np.uint32))
    assert opt.getQuality() == o.surface_mesh_min_quality()
    opt.close()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("fixed", [False, True])
def test_surface_mesh_dmo_parity(oracle, case, fixed):
    """Four iterations: free 2D passes, feature 1D passes (unless fixed),
    nearest-neighbour walk and normal update, bit-exact after each."""
    kind, xyz, f = CASES[case]()
    g = _mesh(kind, xyz, f)
    opt = _opt(kind, g, n_iter=4, fixed_boundary=fixed)
    o = oracle.OracleMesh(kind, xyz, f)
    o.surface_init()
    opt1d = o.surface_get(2)
    for it in range(4):
        opt.doIteration()
        n2, n1 = o.iterate_surface_mesh(fixed_boundary=int(fixed))
        bad = (g.positions(original_order=False) != o.positions()).any(axis=1)
        assert not bad.any(), (case, it, int(bad.sum()), np.nonzero(bad)[0][:10])
        assert opt.getQuality() == o.surface_mesh_min_quality(), it
    st = opt.stats()
    assert st["searches_surface"] == 4 * o.m.nFree
    assert st["searches_feature"] == (0 if fixed else 4 * int(opt1d.sum()))
    assert st["searches_inner"] == 0
    opt.close()


def test_flat_mesh_on_surface_path(oracle):
    """A flat triangle mesh through DMOTriClass takes the surface path (the
    reference CLI's choice for .off files); DMOTriFlatClass keeps the flat path
    and refuses a non-flat mesh."""
    xyz, tris = synth.tri_grid(20, seed=9, jitter=0.4)
    g = DMOMeshTriFactory.create(xyz, tris)
    opt = DMOTriClass(g, n_iter=3)
    o = oracle.OracleMesh("tri", xyz, tris)
    o.surface_init()
    for it in range(3):
        opt.doIteration()
        o.iterate_surface_mesh()
        assert np.array_equal(g.positions(original_order=False), o.positions()), it
        assert opt.getQuality() == o.surface_mesh_min_quality()
    opt.close()
    kind, xs, fs = CASES["tri_field"]()
    g2 = DMOMeshTriFactory.create(xs, fs)
    with pytest.raises(VsdmoError):
        DMOTriFlatClass(g2, n_iter=1)
    with pytest.raises(VsdmoError):
        DMOBatch([g2])


def test_surface_points_on_surface_mesh(oracle):
    kind, xyz, f = CASES["quad_cube"]()
    g = _mesh(kind, xyz, f)
    opt = _opt(kind, g, n_iter=2)
    o = oracle.OracleMesh(kind, xyz, f)
    o.surface_init()
    for _ in range(2):
        opt.doIteration()
        o.iterate_surface_mesh()
    c = g.counts()
    for vid in (0, c["n_free"] // 2, c["n_free"], c["n_surface"] - 1):
        a = opt.getEstimateLocalSurfacePoints(vid, 5, 4)
        b = o.surface_points(vid, 5, 4, 0)
        assert np.array_equal(a, b), vid
        a = opt.getLocalSurfacePoints(vid, 4, 4, 0)
        b = o.surface_points(vid, 4, 4, 1, 0)
        assert np.array_equal(a, b), vid
    opt.close()


@pytest.mark.parametrize("case", ["tri_field", "quad_cube", "tri_sphere"])
def test_laplace_parity(oracle, case):
    kind, xyz, f = CASES[case]()
    g = _mesh(kind, xyz, f)
    o = oracle.OracleMesh(kind, xyz, f)
    laplaceSmoothing(g, 3)
    o.laplace(3)
    assert np.array_equal(g.positions(original_order=False), o.positions())


def test_surface_mesh_at_size(oracle):
    """96 x 96 quads per cube face (55,298 vertices): two iterations
    bit-exact."""
    kind, xyz, f = ("quad",) + synth.cube_surface(96, seed=11, jitter=0.3)
    g = _mesh(kind, xyz, f)
    opt = _opt(kind, g, n_iter=2)
    o = oracle.OracleMesh(kind, xyz, f)
    o.surface_init()
    for it in range(2):
        opt.doIteration()
        o.iterate_surface_mesh()
        bad = (g.positions(original_order=False) != o.positions()).any(axis=1)
        assert not bad.any(), (it, int(bad.sum()))
    opt.close()
