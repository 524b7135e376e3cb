"""Mesh files (host only): TetGen / OFF / MEDIT readers and writers against
hand-written files in the reference formats, and bit-exact round trips."""
import numpy as np
import pytest

from volumesurfacediscretemeshoptimization_amd import VsdmoError, read_mesh_file, synth, write_mesh_file
from volumesurfacediscretemeshoptimization_amd import _lib

TRI, QUAD, TET, HEX = _lib.VSDMO_TRI, _lib.VSDMO_QUAD, _lib.VSDMO_TET, _lib.VSDMO_HEX


def test_tetgen_round_trip(tmp_path):
    xyz, tets = synth.tet_cube(4, seed=3)
    base = str(tmp_path / "cube")
    write_mesh_file(TET, base, xyz, tets)
    x2, t2 = read_mesh_file(TET, base)
    assert np.array_equal(x2.view(np.uint32), xyz.view(np.uint32))
    assert np.array_equal(t2, tets)


def test_tetgen_one_based_comments_attributes(tmp_path):
    (tmp_path / "m.node").write_text(
        "# a comment\n\n5 3 1 1\n"
        "1 0.0 0.0 0.0 7.5 1\n2 1.0 0.0 0.0 7.5 1   # trailing comment\n3 0.0 1.0 0.0 7.5 1\n"
        "4 0.0 0.0 1.0 7.5 1\n5 0.25 0.25 0.25 7.5 0\n")
    (tmp_path / "m.ele").write_text("2 4 0\n1 1 2 3 5\n2 1 2 5 4\n")
    xyz, tets = read_mesh_file(TET, str(tmp_path / "m"))
    assert xyz.shape == (5, 3) and np.allclose(xyz[4], 0.25)
    assert np.array_equal(tets, [[0, 1, 2, 4], [0, 1, 4, 3]])


def test_off_round_trip_and_errors(tmp_path):
    xyz, tris = synth.tri_grid(6, seed=4)
    p = str(tmp_path / "g.off")
    write_mesh_file(TRI, p, xyz, tris)
    assert open(p).readline().strip() == "OFF"
    x2, t2 = read_mesh_file(TRI, p)
    assert np.array_equal(x2.view(np.uint32), xyz.view(np.uint32)) and np.array_equal(t2, tris)
    bad = tmp_path / "bad.off"
    bad.write_text("COFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n")
    with pytest.raises(VsdmoError):
        read_mesh_file(TRI, str(bad))
    quad = tmp_path / "quad.off"
    quad.write_text("OFF\n4 1 0\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n")
    with pytest.raises(VsdmoError):
        read_mesh_file(TRI, str(quad))


def test_quad_off_round_trip_and_errors(tmp_path):
    xyz, quads = synth.cube_surface(5, seed=6)
    p = str(tmp_path / "q.off")
    write_mesh_file(QUAD, p, xyz, quads)
    lines = open(p).read().split("\n")
    assert lines[0] == "OFF" and lines[2 + len(xyz)].startswith("4 ")
    x2, q2 = read_mesh_file(QUAD, p)
    assert np.array_equal(x2.view(np.uint32), xyz.view(np.uint32)) and np.array_equal(q2, quads)
    # a triangle file is not a quad file and vice versa
    t = str(tmp_path / "t.off")
    xt, tt = synth.tri_grid(4, seed=1)
    write_mesh_file(TRI, t, xt, tt)
    with pytest.raises(VsdmoError):
        read_mesh_file(QUAD, t)
    with pytest.raises(VsdmoError):
        read_mesh_file(TRI, p)


def test_medit_reorder_round_trip_and_errors(tmp_path):
    xyz, hexes = synth.hex_cube(2, seed=5, jitter=0.0)
    p = tmp_path / "h.mesh"
    write_mesh_file(HEX, str(p), xyz, hexes)
    lines = p.read_text().splitlines()
    first = lines[lines.index(f"Hexahedra {len(hexes)}") + 1].split()
    # the file's first four vertices are the internal v4..v7, one-based
    assert [int(v) - 1 for v in first[:8]] == list(hexes[0, 4:]) + list(hexes[0, :4])
    x2, h2 = read_mesh_file(HEX, str(p))
    assert np.array_equal(x2, xyz) and np.array_equal(h2, hexes)
    # a Quadrilaterals section is skipped
    txt = p.read_text().replace("Hexahedra", "Quadrilaterals 1\n1 2 3 4 0\nHexahedra", 1)
    (tmp_path / "q.mesh").write_text(txt)
    x3, h3 = read_mesh_file(HEX, str(tmp_path / "q.mesh"))
    assert np.array_equal(h3, hexes)
    # an unreferenced vertex and an index out of range are errors
    unused = "MeshVersionFormatted 1\nDimension 3\nVertices 9\n" + "0 0 0 0\n" * 9 + \
        "Hexahedra 1\n1 2 3 4 5 6 7 8 0\nEnd\n"
    (tmp_path / "u.mesh").write_text(unused)
    with pytest.raises(VsdmoError):
        read_mesh_file(HEX, str(tmp_path / "u.mesh"))
    oor = unused.replace("Vertices 9", "Vertices 8").replace("0 0 0 0\n" * 9, "0 0 0 0\n" * 8).replace(
        "1 2 3 4 5 6 7 8 0", "1 2 3 4 5 6 7 9 0")
    (tmp_path / "o.mesh").write_text(oor)
    with pytest.raises(VsdmoError):
        read_mesh_file(HEX, str(tmp_path / "o.mesh"))
    (tmp_path / "t.mesh").write_text("MeshVersionFormatted 1\nBogus 3\n")
    with pytest.raises(VsdmoError):
        read_mesh_file(HEX, str(tmp_path / "t.mesh"))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(VsdmoError):
        read_mesh_file(TET, str(tmp_path / "nothing"))
