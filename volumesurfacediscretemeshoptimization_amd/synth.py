"""Seeded synthetic meshes for the DMO workloads.

All generators are deterministic functions of their seed.  Random numbers come
from a counter-based generator: draw ``idx`` of stream ``seed`` is the
SplitMix64 finaliser of ``seed' + idx * 0x9E3779B97F4A7C15`` and
``u = (z >> 40) * 2**-24`` in [0, 1).  Vertex ``v`` uses draws ``3v+0, 3v+1,
3v+2`` for x, y, z, so any vertex's jitter can be recomputed in isolation.
Coordinates are computed in float64 and rounded once to float32.

Workloads (sizes as used by the benchmark and the tests):

* ``tri_grid(64)``     2D triangle grid, 64x64 vertices on [0,1]^2, h = 1/63,
  cells split on the (v00, v11) diagonal, interior jitter +-0.4h, z = 0.
* ``tet_cube(40)``     unit cube, 41^3 vertices, 40^3 cubes, each cut into the
  6 Kuhn tets [c000, c000+e_a, c000+e_a+e_b, c111] (384,000 tets), interior
  jitter +-0.35h, boundary fixed.
* ``hex_cube(48)``     unit cube, 49^3 vertices, 48^3 hexes in the internal
  vertex order of the optimiser (top face z=1 first, then bottom), interior
  jitter +-0.3h, face vertices jittered +-0.3h in both tangent directions,
  edge vertices along their edge, corners fixed.
* ``batch_mix(256)``   256 meshes, sizes log-uniform in [500, 50000]; even
  index: tri grid with m = round(sqrt n) vertices per side; odd: tet cube with
  m = round(cbrt n) vertices per side; jitter +-0.35h.
* ``cube_surface(n)``  surface of the unit cube, n x n quads per face
  (outward), or each quad (a, b, c, d) split into (a, b, c), (a, c, d); face
  vertices jittered +-jitter*h in both tangent directions, edge vertices
  along their edge, corners fixed.  ``sphere=True`` maps the jittered cube
  [-1, 1]^3 surface radially onto the unit sphere.
* ``height_field(n)``  open surface z = 0.2 sin(1.5 pi x) cos(1.25 pi y) +
  0.1 x y over an n x n vertex grid of [0, 1]^2 (quads, or the tri_grid
  split); interior (x, y) jittered +-jitter*h, boundary vertices along their
  side, corners fixed.
"""
from __future__ import annotations

import itertools

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_SEEDMIX = np.uint64(0x5851F42D4C957F2D)


def uniform(seed: int, idx: np.ndarray) -> np.ndarray:
    """Counter-based uniform draws in [0, 1) (float64, 24-bit resolution)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ _SEEDMIX) + np.asarray(idx, dtype=np.uint64) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def _jitter(seed: int, nv: int) -> np.ndarray:
    """(nv, 3) array of 2u-1 in [-1, 1)."""
    u = uniform(seed, np.arange(3 * nv, dtype=np.uint64)).reshape(nv, 3)
    return 2.0 * u - 1.0


def tri_grid(n_side: int = 64, seed: int = 0, jitter: float = 0.4):
    """2D flat triangle grid: returns (xyz float32 (nv,3), tris int32 (nt,3))."""
    m = n_side
    h = 1.0 / (m - 1)
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    x = (i * h).ravel().astype(np.float64)
    y = (j * h).ravel().astype(np.float64)
    inner = ((i > 0) & (i < m - 1) & (j > 0) & (j < m - 1)).ravel()
    r = _jitter(seed, m * m)
    x = x + np.where(inner, r[:, 0] * jitter * h, 0.0)
    y = y + np.where(inner, r[:, 1] * jitter * h, 0.0)
    xyz = np.stack([x, y, np.zeros_like(x)], axis=1).astype(np.float32)
    cj, ci = np.meshgrid(np.arange(m - 1), np.arange(m - 1), indexing="ij")
    v00 = (cj * m + ci).ravel()
    v10 = v00 + 1
    v01 = v00 + m
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3).astype(np.int32)
    return xyz, tris


def _grid3(m: int):
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def tet_cube(n_cells: int = 40, seed: int = 0, jitter: float = 0.35):
    """Kuhn-split unit cube: returns (xyz float32 (nv,3), tets int32 (nt,4))."""
    m = n_cells + 1
    h = 1.0 / n_cells
    i, j, k = _grid3(m)
    pos = np.stack([i, j, k], axis=1).astype(np.float64) * h
    inner = (i > 0) & (i < m - 1) & (j > 0) & (j < m - 1) & (k > 0) & (k < m - 1)
    r = _jitter(seed, m ** 3)
    pos = pos + np.where(inner[:, None], r * jitter * h, 0.0)
    xyz = pos.astype(np.float32)
    ck, cj, ci = np.meshgrid(np.arange(n_cells), np.arange(n_cells), np.arange(n_cells), indexing="ij")
    base = (ci + m * (cj + m * ck)).ravel()
    step = np.array([1, m, m * m])
    c111 = base + step.sum()
    tets = []
    for perm in itertools.permutations(range(3)):
        a, b, _ = perm
        v1 = base + step[a]
        v2 = v1 + step[b]
        e = [np.eye(3)[a], np.eye(3)[a] + np.eye(3)[b], np.ones(3)]
        det = np.linalg.det(np.stack(e))
        if det > 0:
            tets.append(np.stack([base, v1, v2, c111], axis=1))
        else:
            tets.append(np.stack([base, v1, c111, v2], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return xyz, tets


def hex_cube(n_cells: int = 48, seed: int = 0, jitter: float = 0.3):
    """Unit cube hex grid in the optimiser's internal vertex order (top first)."""
    m = n_cells + 1
    h = 1.0 / n_cells
    i, j, k = _grid3(m)
    pos = np.stack([i, j, k], axis=1).astype(np.float64) * h
    on = np.stack([(i == 0) | (i == m - 1), (j == 0) | (j == m - 1), (k == 0) | (k == m - 1)], axis=1)
    nb = on.sum(axis=1)
    r = _jitter(seed, m ** 3)
    # interior: all axes; face (one bounding axis): the two tangent axes; edge
    # (two bounding axes): the one free axis; corner: none
    move = ~on
    pos = pos + np.where(move, r * jitter * h, 0.0)
    del nb
    xyz = pos.astype(np.float32)
    ck, cj, ci = np.meshgrid(np.arange(n_cells), np.arange(n_cells), np.arange(n_cells), indexing="ij")
    b00 = (ci + m * (cj + m * ck)).ravel()
    dx, dy, dz = 1, m, m * m
    bot = [b00, b00 + dx, b00 + dx + dy, b00 + dy]
    top = [v + dz for v in bot]
    hexes = np.stack(top + bot, axis=1).astype(np.int32)
    return xyz, hexes


def batch_mix(n_meshes: int = 256, seed: int = 0, lo: int = 500, hi: int = 50000, jitter: float = 0.35):
    """List of (kind, xyz, cells) with kind 'tri' or 'tet'."""
    u = uniform(seed ^ 0xB47C, np.arange(n_meshes, dtype=np.uint64))
    sizes = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    out = []
    for b in range(n_meshes):
        n = float(sizes[b])
        if b % 2 == 0:
            m = max(3, int(round(np.sqrt(n))))
            xyz, cells = tri_grid(m, seed=seed + 1000 + b, jitter=jitter)
            out.append(("tri", xyz, cells))
        else:
            m = max(3, int(round(np.cbrt(n))))
            xyz, cells = tet_cube(m - 1, seed=seed + 1000 + b, jitter=jitter)
            out.append(("tet", xyz, cells))
    return out


def cube_surface(n: int = 16, seed: int = 0, jitter: float = 0.3, quads: bool = True, sphere: bool = False):
    """Closed cube (or cube-sphere) surface: (xyz float32, faces int32) with
    quads (nf, 4) or triangles (nf, 3), consistently oriented outward."""
    m = n + 1
    i, j, k = _grid3(m)
    lat = np.stack([i, j, k], axis=1)
    on = (lat == 0) | (lat == n)
    onb = on.any(axis=1)
    gid = np.full(m ** 3, -1, dtype=np.int64)
    gid[onb] = np.arange(int(onb.sum()))
    r = _jitter(seed, m ** 3)
    h = 1.0 / n
    pos = lat.astype(np.float64) * h + np.where(~on, r * jitter * h, 0.0)
    pos = pos[onb]
    if sphere:
        pos = 2.0 * pos - 1.0
        pos = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    U, V = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    U, V = U.ravel(), V.ravel()
    faces = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for side in (0, n):
            def corner(du, dv):
                L = np.zeros((n * n, 3), dtype=np.int64)
                L[:, a] = side
                L[:, b] = U + du
                L[:, c] = V + dv
                return gid[L[:, 0] + m * (L[:, 1] + m * L[:, 2])]
            q = np.stack([corner(0, 0), corner(1, 0), corner(1, 1), corner(0, 1)], axis=1)
            if side == 0:
                q = q[:, ::-1]
            faces.append(q)
    q = np.concatenate(faces)
    if not quads:
        q = np.stack([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1, 3)
    return pos.astype(np.float32), q.astype(np.int32)


def height_field(n_side: int = 32, seed: int = 0, jitter: float = 0.3, quads: bool = True):
    """Open height-field surface with a boundary: (xyz float32, faces int32)."""
    m = n_side
    h = 1.0 / (m - 1)
    j, i = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    i, j = i.ravel(), j.ravel()
    bi = (i == 0) | (i == m - 1)
    bj = (j == 0) | (j == m - 1)
    r = _jitter(seed, m * m)
    x = i * h + np.where(~bi, r[:, 0] * jitter * h, 0.0)
    y = j * h + np.where(~bj, r[:, 1] * jitter * h, 0.0)
    z = 0.2 * np.sin(1.5 * np.pi * x) * np.cos(1.25 * np.pi * y) + 0.1 * x * y
    xyz = np.stack([x, y, z], axis=1).astype(np.float32)
    cj, ci = np.meshgrid(np.arange(m - 1), np.arange(m - 1), indexing="ij")
    v00 = (cj * m + ci).ravel()
    v10, v01 = v00 + 1, v00 + m
    v11 = v01 + 1
    if quads:
        cells = np.stack([v00, v10, v11, v01], axis=1)
    else:
        cells = np.stack([np.stack([v00, v10, v11], axis=1), np.stack([v00, v11, v01], axis=1)], axis=1).reshape(-1, 3)
    return xyz, cells.astype(np.int32)
