"""Pins the oracle's searches against the independent FP64 restatement of the
reference (tests/ref64.py), CPU only.  Each test names the reference code it
pins; tests/test_oracle_mutations.py shows that each check rejects a planted
one-line mistake."""
import pytest

import oracle_pins as P


def test_tet_search_matches_restatement(oracle):
    """k_optimizeHierarchicalInnerTet (dmo/DMOTetImplCub.h:20-137) on the 125
    interior vertices of a jittered 6^3 tet cube"""
    r = P.check_volume_search(oracle, "tet", n=6, seed=5, jitter=0.35)
    assert r["vertices"] == 125 and r["agree"] >= 120


def test_tet_exact_tie_lowest_index(oracle):
    """cub::ArgMax keeps the lower thread and each thread the first of equal
    candidates: an exact tie goes to the lowest candidate index"""
    P.check_tet_tie(oracle)


def test_hex_metric_principal_axes(oracle):
    """shapeMetricHex (dmo/DMOQualityHex.h:54-71) on hexes whose minimum is the
    principal-axes sub-tet, and on random distorted hexes"""
    P.check_hex_metric(oracle)


def test_hex_search_matches_restatement(oracle):
    """k_optimizeHierarchicalInnerHex (dmo/DMOHexImplCub.h:20-137): the window
    is the per-axis minimum extent"""
    r = P.check_volume_search(oracle, "hex", n=5, seed=7, jitter=0.2, min_clean=0.75)
    assert r["vertices"] == 64


def test_flat_search_matches_restatement(oracle):
    """k_optimizeHierarchical2DFlat (dmo/DMOTriFlatImplCub.h:23-152)"""
    P.check_flat_search(oracle)


def test_flat_exact_tie_takes_b(oracle):
    """an exact A/B tie goes to B in the flat search"""
    P.check_flat_tie(oracle)


def test_flat_last_level_write_rule(oracle):
    """the flat search writes only when the last level's maximum beats the
    start, dropping an earlier level's gain"""
    P.check_flat_write_rule(oracle)


def test_surface2d_search_tet(oracle):
    """k_optimizeHierarchical2DTet (dmo/DMOTetImplCub.h:141-322): Euclidean
    max window, A unless q_B > q_A, the last-level write check"""
    P.check_surface_search(oracle, "tet", 2)


def test_surface2d_search_hex(oracle):
    """k_optimizeHierarchical2DHex (dmo/DMOHexImplCub.h:141-318): Euclidean
    min window"""
    P.check_surface_search(oracle, "hex", 2)


def test_surface2d_exact_tie_takes_a(oracle):
    P.check_surface_tie(oracle)


def test_feature1d_search_tet(oracle):
    """k_optimizeHierarchical1DTet (dmo/DMOTetImplCub.h:326-494)"""
    P.check_surface_search(oracle, "tet", 1)


def test_feature1d_search_hex(oracle):
    """k_optimizeHierarchical1DHex (dmo/DMOHexImplCub.h:320-489)"""
    P.check_surface_search(oracle, "hex", 1)


def test_projection_blend_and_sectors(oracle):
    """localSurfaceEstimationWithFeature (dmo/SurfaceEstimation.h:122-197),
    including the feature neighbours' sector lookup"""
    assert P.check_projection(oracle) <= 1e-6


def test_nearest_neighbour_walk(oracle):
    """k_updateNearestNeighbor (dmo/DMOCommon.cu:11-43) against the exhaustive
    nearest initial position of the vertex's class"""
    assert P.check_nn_walk(oracle) >= 0.95


def test_learned_window_matches_restatement(oracle):
    """the learned proposal's window search (this build's own; its
    specification restated in tests/ref64.py) at the shipped windows, tet
    (1 level, 0.025) and hex (1, 2.0), the specified 1/4 window, and two
    levels"""
    for kind, depth, frac in (("tet", 1, 0.025), ("tet", 1, 0.25), ("tet", 2, 0.5), ("hex", 1, 2.0), ("hex", 2, 0.5)):
        st = P.check_learned_window(oracle, kind, depth, frac)
        assert st[1] > 0 and st[-1] > 0, (kind, depth, frac, st)
        if kind == "hex":
            assert st[0] > 0, st          # vertices kept in place by the positive-tie / pinned rules


@pytest.mark.parametrize("kind", ["tri", "quad"])
@pytest.mark.parametrize("dim,shape", [(2, "sphere"), (2, "height"), (1, "height")])
def test_fan_search_matches_restatement(oracle, kind, dim, shape):
    """k_optimizeHierarchical2D / 1D of triangle and quad surface meshes
    (dmo/DMOTriImplCub.h:25-378, dmo/DMOQuadImplCub.h:23-402) on a
    cube-sphere and an open height field"""
    P.check_fan_search(oracle, kind, dim, shape)


def test_flat_boundary_1d_matches_restatement(oracle):
    """k_optimizeHierarchical1DFlat (dmo/DMOTriFlatImplCub.h:158-333)"""
    P.check_flat1d_search(oracle)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_feature_sectors_follow_the_specification(oracle, kind):
    P.check_sectors(oracle, kind)


def test_hex_corner8_metric(oracle):
    """the eight-corner hex metric (vsdmo_hex_metric CORNER8): shapeMetricHex
    (dmo/DMOQualityHex.h:54-71) without the principal-axes sub-tet, against
    the restatement; it scores the twisted hexes above the nine-tet metric"""
    P.check_hex_corner8(oracle)


def test_hex_corner8_search_matches_restatement(oracle):
    """the inner hex search (dmo/DMOHexImplCub.h:20-137) on the eight-corner
    metric: a jittered cube, and single twisted hexes whose minimum is the
    principal-axes sub-tet"""
    r = P.check_volume_search(oracle, "hex", n=5, seed=7, jitter=0.2, min_clean=0.75, corner8=True)
    assert r["vertices"] == 64
    P.check_hex_corner8_rings(oracle)


@pytest.mark.parametrize("dim", [2, 1])
def test_hex_corner8_surface_search(oracle, dim):
    """the surface 2D / feature 1D hex searches (dmo/DMOHexImplCub.h:141-489)
    on the eight-corner metric"""
    P.check_surface_search(oracle, "hex", dim, corner8=True)


def test_hex_corner8_learned_window(oracle):
    st = P.check_learned_window(oracle, "hex", 1, 2.0, corner8=True)
    assert st[1] > 0 and st[-1] > 0, st


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_ring2_fits(oracle, kind):
    """quadric fits over {v} + one-ring + 2-ring (vsdmo_params.surface_ring
    2; the reference fits the one-ring, dmo/surfaceP/Surface.cu:11-105):
    point sets from the face graph and FP64 least squares"""
    P.check_ring2_fits(oracle, kind)


@pytest.mark.parametrize("kind,dim", [("tet", 2), ("hex", 2), ("hex", 1)])
def test_surface_search_on_ring2_fits(oracle, kind, dim):
    """the surface searches on 2-ring fits (the restatement reads the fits
    from the oracle, pinned by test_ring2_fits)"""
    P.check_surface_search(oracle, kind, dim, ring=2)
