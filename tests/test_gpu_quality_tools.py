"""GPU tests of the quality histogram and of the colour-class check.

Histogram (k_qualityHistogram, dmo/DMOTet.cu:31-43, read by
getQualityHistogram / displayQualityGPU, dmo/DMOBase.h:61-87): element
quality q goes to bin int(min(0.9999, q) * n_cols) in FP32; the bins must
equal the counts of the oracle's element qualities (bit-exact with the
GPU's).  An inverted element's negative quality goes to bin 0 here; the
reference converts the negative product to size_t, which is undefined.

Colour classes (k_checkColoring, mesh/MeshGPUCommon.cu:1470-1680): every
class the mesh build produces is an independent set of the element graph;
a partition that merges two classes is rejected with VSDMO_TOPOLOGY.
"""
import numpy as np
import pytest

from volumesurfacediscretemeshoptimization_amd import (DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                       DMOMeshTriFactory, DMOTetClass, VsdmoError, synth)
from volumesurfacediscretemeshoptimization_amd import _lib

pytestmark = pytest.mark.gpu


def _bins(q, n_cols):
    q = np.minimum(np.float32(0.9999), q.astype(np.float32))
    idx = np.trunc(q * np.float32(n_cols)).astype(np.int64)
    return np.bincount(np.maximum(idx, 0), minlength=n_cols)[:n_cols]


@pytest.mark.parametrize("kind,n_cols", [("tet", 10), ("tet", 7), ("hex", 10)])
def test_histogram_matches_oracle_bins(oracle, kind, n_cols):
    if kind == "tet":
        xyz, cells = synth.tet_cube(10, seed=3)
        g = DMOMeshTetFactory.create(xyz, cells)
        opt = DMOTetClass(g, n_iter=2, fixed_boundary=True)
    else:
        xyz, cells = synth.hex_cube(8, seed=3, jitter=0.45)       # inverted hexes: quality 0
        g = DMOMeshHexFactory.create(xyz, cells)
        opt = DMOHexClass(g, n_iter=2, fixed_boundary=True)
    o = oracle.OracleMesh(kind, xyz, cells)
    for it in range(3):
        h = opt.getQualityHistogram(n_cols)
        oq = o.element_qualities()
        assert np.array_equal(opt.getElementQualities(), oq)
        assert np.array_equal(h, _bins(oq, n_cols)), (it, h, _bins(oq, n_cols))
        assert h.sum() == len(cells)
        if it < 2:
            opt.doIteration()
            o.iterate()
    opt.displayQualityGPU(n_cols)
    opt.close()
    g.close()


def test_histogram_inverted_elements_bin_zero(oracle):
    """a jittered tet cube with inverted tets (negative quality): all of them in bin 0"""
    xyz, cells = synth.tet_cube(8, seed=9, jitter=0.9)
    g = DMOMeshTetFactory.create(xyz, cells)
    opt = DMOTetClass(g, n_iter=1, fixed_boundary=True)
    q = opt.getElementQualities()
    assert (q < 0).sum() > 0
    h = opt.getQualityHistogram(10)
    assert np.array_equal(h, _bins(q, 10)) and h[0] >= (q < 0).sum()
    opt.close()
    g.close()


@pytest.mark.parametrize("kind", ["tet", "hex", "tri"])
def test_colour_classes_checked(kind):
    if kind == "tet":
        g = DMOMeshTetFactory.create(*synth.tet_cube(6, seed=1))
    elif kind == "hex":
        g = DMOMeshHexFactory.create(*synth.hex_cube(6, seed=1, jitter=0.2))
    else:
        g = DMOMeshTriFactory.create(*synth.tri_grid(12, seed=1))
    whiches = ["inner", "free", "feature"] if kind != "tri" else ["free", "feature"]
    for which in whiches:
        off = g.color_offsets(which)
        if off[-1] == off[0]:
            continue
        g.check_coloring(off)                                   # the build's classes pass
        g.check_coloring(np.arange(off[0], off[-1] + 1))        # one vertex per class passes
        if len(off) > 2:
            merged = np.delete(off, 1)                           # classes 0 and 1 as one
            with pytest.raises(VsdmoError) as e:
                g.check_coloring(merged)
            assert e.value.status == _lib.VSDMO_TOPOLOGY and "class 0" in str(e.value)
    with pytest.raises(VsdmoError) as e:
        g.check_coloring([0, g.nVertices + 1])
    assert e.value.status == _lib.VSDMO_INVALID_ARG
    g.close()
