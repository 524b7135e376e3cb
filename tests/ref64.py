"""FP64 restatement of the reference's per-vertex DMO searches (TEST
INFRASTRUCTURE), written from the reference sources to pin the oracle.

It shares no code with the oracle.  Each function follows the reference
file:line it cites literally: block layouts (threads, candidates per thread,
cub::ArgMax keeping the lower thread on equal values), the per-thread
running maxima, the move and write conditions, the windows and candidate
grids.  Grids, windows and local coordinates are computed in FP32 with one
rounding per operation, as the reference computes them; every quality and
every projection is computed in FP64.  Two FP64 values closer than 1e-12
relative count as equal (the stand-in for an exact tie).  Each search also
returns its smallest decision margin: the tests compare the oracle bit for
bit wherever that margin is clear of FP32 rounding.

Paths are relative to the reference repository, VSDMO/src/ shortened.
"""
from __future__ import annotations

import numpy as np

F = np.float32
EPS = F(1.19209e-07)          # EPSILON, dmo/SurfaceCommon.h:8
TIE = 1e-12


def eq(a: float, b: float) -> bool:
    return abs(a - b) <= TIE * max(abs(a), abs(b), 1e-300)


def gt(a: float, b: float) -> bool:
    """a > b outside the tie band"""
    return a > b and not eq(a, b)


def rel_gap(a: float, b: float) -> float:
    """relative gap of a decision between a and b; an exact tie (decided by
    the reference's tie rules, not by rounding) does not narrow the margin"""
    if eq(a, b):
        return np.inf
    return abs(a - b) / max(abs(a), abs(b), 1e-30)


def _dot(a, b):
    return np.einsum("...i,...i->...", a, b)


# ------------------------------------------------------------ element metrics
def q_tet(P):
    """meanRatioMetricTet (dmo/DMOQualityTet.h:119-143): signed volume when
    negative, else 12 cbrt(9) cbrt(V^2) / sum of the six squared edges.
    P[..., 4, 3]."""
    P = np.asarray(P, np.float64)
    e = [P[..., 1, :] - P[..., 0, :], P[..., 2, :] - P[..., 0, :], P[..., 3, :] - P[..., 0, :],
         P[..., 2, :] - P[..., 1, :], P[..., 3, :] - P[..., 1, :], P[..., 3, :] - P[..., 2, :]]
    vol = (1.0 / 6.0) * _dot(e[0], np.cross(e[1], e[2]))
    l = sum(_dot(x, x) for x in e)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = 12.0 * np.cbrt(9.0) * np.cbrt(vol * vol) / l
    return np.where(vol < 0.0, vol, q)


def hex_subtets(P):
    """The nine sub-tets of shapeMetricHex (dmo/DMOQualityHex.h:54-71):
    hex_edges / hex_principal_axes on (p4..p7, p0..p3) and hex_subtets
    (dmo/DMOQualityUtil.h:108-149).  Returns (det, |J|^2), each [..., 9]."""
    P = np.asarray(P, np.float64)
    q = [P[..., k, :] for k in (4, 5, 6, 7, 0, 1, 2, 3)]
    L = [q[1] - q[0], q[2] - q[1], q[3] - q[2], q[3] - q[0], q[4] - q[0], q[5] - q[1], q[6] - q[2], q[7] - q[3],
         q[5] - q[4], q[6] - q[5], q[7] - q[6], q[7] - q[4]]
    X = [(q[1] - q[0]) + (q[2] - q[3]) + (q[5] - q[4]) + (q[6] - q[7]),
         (q[3] - q[0]) + (q[2] - q[1]) + (q[7] - q[4]) + (q[6] - q[5]),
         (q[4] - q[0]) + (q[5] - q[1]) + (q[6] - q[2]) + (q[7] - q[3])]
    T = [(L[0], L[3], L[4]), (L[1], -L[0], L[5]), (L[2], -L[1], L[6]), (-L[3], -L[2], L[7]),
         (L[11], L[8], -L[4]), (-L[8], L[9], -L[5]), (-L[9], L[10], -L[6]), (-L[10], -L[11], -L[7]),
         (X[0], X[1], X[2])]
    det = np.stack([_dot(t0, np.cross(t1, t2)) for t0, t1, t2 in T], axis=-1)
    jn = np.stack([_dot(t0, t0) + _dot(t1, t1) + _dot(t2, t2) for t0, t1, t2 in T], axis=-1)
    return det, jn


def q_hex(P, corner8=False):
    """shapeMetricHex (dmo/DMOQualityHex.h:54-71): 0 when any sub-tet has
    det < EPSILON or |J|^2 < EPSILON, else 3 min cbrt(det^2) / |J|^2.
    corner8: the same over the eight corner sub-tets (the first eight of
    hex_subtets) without the principal-axes tet."""
    det, jn = hex_subtets(P)
    if corner8:
        det, jn = det[..., :8], jn[..., :8]
    bad = np.any((det < float(EPS)) | (jn < float(EPS)), axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.min(np.cbrt(det * det) / jn, axis=-1)
    return np.where(bad, 0.0, 3.0 * v)


def q_tri2d(P):
    """meanRatioMetricTri2D (dmo/DMOQualityTri.h:243-257): the signed area
    e0 x e1 when negative, else 2 sqrt(3) area / sum of squared edges."""
    P = np.asarray(P, np.float64)[..., :2]
    e = [P[..., 1, :] - P[..., 0, :], P[..., 2, :] - P[..., 1, :], P[..., 0, :] - P[..., 2, :]]
    l = sum(x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1] for x in e)
    area = e[0][..., 0] * e[1][..., 1] - e[0][..., 1] * e[1][..., 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        q = 2.0 * np.sqrt(3.0) * area / l
    return np.where(area < 0.0, area, q)


def ring_quality(kind: str, ring, p, corner8=False):
    """Ring quality at candidate points p[..., 3]: the minimum over the ring
    elements with p as vertex 0 (qualityShapeMRMTet, DMOQualityTet.h:146-162;
    qualityShapeMRMHex, DMOQualityHex.h:75-95; the closed 2D fan of
    qualityShapeMRMTri2D, DMOQualityTri.h:290-305)."""
    p = np.asarray(p, np.float64)
    ring = np.asarray(ring, np.float64)
    lead = p.shape[:-1]
    pp = p.reshape(-1, 1, 1, 3)
    if kind == "tri":
        r = ring.reshape(-1, 3)
        els = np.stack([r, np.roll(r, -1, axis=0)], axis=1)           # (r_k, r_k+1), r_n = r_0
        P = np.concatenate([np.broadcast_to(pp, (pp.shape[0], len(els), 1, 3)),
                            np.broadcast_to(els, (pp.shape[0],) + els.shape)], axis=2)
        return q_tri2d(P).min(axis=1).reshape(lead)
    per = 3 if kind == "tet" else 7
    els = ring.reshape(-1, per, 3)
    P = np.concatenate([np.broadcast_to(pp, (pp.shape[0], len(els), 1, 3)),
                        np.broadcast_to(els, (pp.shape[0],) + els.shape)], axis=2)
    q = q_tet(P) if kind == "tet" else q_hex(P, corner8)
    return q.min(axis=1).reshape(lead)


# ------------------------------------------------------------- block helpers
def _argmax_threads(vals):
    """cub::ArgMax over (thread, value): the largest value, the lower
    thread among equal values."""
    key = 0
    for t in range(1, len(vals)):
        if gt(vals[t], vals[key]):
            key = t
    return key


def _top_gap(vals):
    """relative gap between the largest value and the next distinct one"""
    v = np.sort(np.asarray(vals, np.float64))[::-1]
    for x in v[1:]:
        if not eq(x, v[0]):
            return rel_gap(v[0], x)
    return np.inf


def _grid(aff, i, nq, gmin, gmax):
    """affineFactor * (i * gridMin + (NQ - 1 - i) * gridMax) in FP32"""
    i = np.asarray(i)
    return F(aff) * (i.astype(F) * F(gmin) + (F(nq - 1) - i.astype(F)) * F(gmax))


class Result:
    def __init__(self, pos, written, margin, extra=None):
        self.pos, self.written, self.margin, self.extra = pos, written, margin, extra or {}


# ------------------------------------------------------------ inner searches
def search_volume(kind: str, x0, ring, nq=8, depth=3, grid_scale=0.5, block=128, corner8=False) -> Result:
    """k_optimizeHierarchicalInnerTet / Hex (dmo/DMOTetImplCub.h:20-137,
    dmo/DMOHexImplCub.h:20-137), launched with 128 threads
    (dmo/DMOTet.cu:134, dmo/DMOHex.cu:146): thread t scans candidates
    [t*ppt, (t+1)*ppt) keeping a running maximum over all levels; the window
    half-width per axis is the largest (tet) or smallest (hex) |ring - x0|."""
    x0 = np.asarray(x0, F).reshape(3)
    pts = np.asarray(ring, F).reshape(-1, 3)
    d = np.abs(pts - x0)
    ext = d.min(axis=0) if kind == "hex" else d.max(axis=0)
    aff = F(1.0) / F(nq - 1)
    ppt = nq ** 3 // block
    idx = np.arange(nq ** 3)
    ii, jj, kk = idx % nq, (idx // nq) % nq, idx // (nq * nq)
    cur = x0.copy()
    ds = F(grid_scale)
    qmax = np.full(block, -np.float64(np.finfo(F).max))
    margin = np.inf
    levels = []
    key = 0
    for _ in range(depth):
        gmax = cur + ds * ext
        gmin = cur - ds * ext
        cand = np.stack([_grid(aff, ii, nq, gmin[0], gmax[0]), _grid(aff, jj, nq, gmin[1], gmax[1]),
                         _grid(aff, kk, nq, gmin[2], gmax[2])], axis=1)
        q = ring_quality(kind, pts, cand, corner8)
        qlast = float(ring_quality(kind, pts, cur, corner8))
        pmax = [None] * block
        for t in range(block):
            for c in range(t * ppt, (t + 1) * ppt):
                if gt(q[c], qmax[t]):
                    qmax[t] = q[c]
                    pmax[t] = c
        key = _argmax_threads(qmax)
        margin = min(margin, _top_gap(q), rel_gap(qmax[key], qlast))
        moved = gt(qmax[key], qlast)
        if moved:
            cur = cand[pmax[key]].copy()
        levels.append((int(pmax[key]) if pmax[key] is not None else -1, moved))
        ds = ds * (F(2.0) / F(nq - 1))
    qold = float(ring_quality(kind, pts, x0, corner8))
    margin = min(margin, rel_gap(qmax[key], qold))
    written = gt(qmax[key], qold)
    return Result(cur if written else x0, written, margin, {"levels": levels})


def search_flat(x0, fan, nq=8, depth=3, grid_scale=0.5, shrink=None) -> Result:
    """k_optimizeHierarchical2DFlat (dmo/DMOTriFlatImplCub.h:23-152) with
    nq^2/2 threads: thread t holds A = (t/nq, t%nq) and B = (t/nq + nq/2,
    t%nq) and takes B unless q_A > q_B; the block maximum is per level; the
    position is written iff q(old) < the LAST level's maximum.  The closed fan
    of the halfedge targets is the ring; the window half-width per axis is
    the largest |ring - x0|.  shrink: the per-level window factor (None: the
    reference's 2/(nq - 1))."""
    x0 = np.asarray(x0, F).reshape(3)
    r = np.asarray(fan, F).reshape(-1, 3)
    ext = np.abs(r - x0).max(axis=0)
    aff = F(1.0) / F(nq - 1)
    T, half = nq * nq // 2, nq // 2
    t = np.arange(T)
    i1, j1 = t // nq, t % nq
    i2 = i1 + half
    cur = x0.copy()
    ds = F(grid_scale)
    shr = (F(2.0) / F(nq - 1)) if shrink is None else F(shrink)
    margin = np.inf
    improved_early = False
    qm = None
    key = 0
    for lev in range(depth):
        gmax = cur + ds * ext
        gmin = cur - ds * ext
        A = np.stack([_grid(aff, i1, nq, gmin[0], gmax[0]), _grid(aff, j1, nq, gmin[1], gmax[1]), np.zeros(T, F)], 1)
        B = np.stack([_grid(aff, i2, nq, gmin[0], gmax[0]), A[:, 1], np.zeros(T, F)], 1)
        qa = ring_quality("tri", r, A)
        qb = ring_quality("tri", r, B)
        qlast = float(ring_quality("tri", r, cur))
        qm = np.where([gt(b, a) for a, b in zip(qa, qb)], qb, qa)
        key = _argmax_threads(qm)
        margin = min(margin, _top_gap(np.concatenate([qa, qb])), rel_gap(qm[key], qlast))
        if gt(qm[key], qlast):
            cur = (A[key] if gt(qa[key], qb[key]) else B[key]).copy()
            cur[2] = x0[2]
            if lev < depth - 1:
                improved_early = True
        ds = ds * shr
    qold = float(ring_quality("tri", r, x0))
    margin = min(margin, rel_gap(qm[key], qold))
    written = gt(qm[key], qold)
    return Result(cur if written else x0, written, margin,
                  {"dropped_gain": bool(improved_early and not written)})


# ----------------------------------------------------------- surface vertices
class SurfaceState:
    """What the surface searches read, taken from an oracle mesh after its
    surface init: positions, halfedge fans, feature flags, the local fits
    (pinned on their own), the sector table and fits, the 1D curves, the
    initial positions and the nearest-neighbour table."""

    def __init__(self, kind, pos, he_row, he_target, is_feature, n_free, ls, ls1, table, frow, fits, pinit, nn,
                 opt1d, ring_row, ring_val, he_next=None, is_b1d=None, vnorm=None):
        self.kind = kind
        self.he_next, self.is_b1d = he_next, is_b1d
        self.vnorm = None if vnorm is None else np.asarray(vnorm, np.float64)
        self.pos = np.asarray(pos, F)
        self.he_row, self.he_target = he_row, he_target
        self.is_feature, self.n_free = is_feature, n_free
        self.ls = np.asarray(ls, np.float64)
        self.ls1 = np.asarray(ls1, np.float64)
        self.table, self.frow, self.fits = table, frow, np.asarray(fits, np.float64)
        self.pinit = np.asarray(pinit, np.float64)
        self.nn, self.opt1d = nn, opt1d
        self.ring_row, self.ring_val = ring_row, ring_val
        self.corner8 = False      # hex ring scores over the eight corner sub-tets

    def fan(self, v):
        return self.he_target[self.he_row[v]:self.he_row[v + 1]]

    def vring(self, v):
        return self.pos[self.ring_val[self.ring_row[v]:self.ring_row[v + 1]].ravel()]


def _poly(s, u, v):
    return s[12] * u * u + s[13] * u * v + s[14] * v * v


def _l2w(s, u, v, w):
    return s[0:3] + u * s[3:6] + v * s[6:9] + w * s[9:12]


def project2d(S: SurfaceState, vid: int, u: float, v: float) -> np.ndarray:
    """localSurfaceEstimationWithFeature (dmo/SurfaceEstimation.h:122-197):
    the point (u, v, w(u, v)) of the NN's fit, and its projection onto the fit
    of each neighbour's NN -- for a feature NN the fit of the sector whose
    halfedge target lies nearest to the vertex -- blended with the weights
    1 / (|p_init(NN) - p|^2 + EPSILON).  In FP64."""
    eps = float(EPS)
    L = S.ls[S.nn[vid]]
    pw = _l2w(L, u, v, _poly(L, u, v))
    pts = [pw]
    d = S.pinit[S.nn[vid]] - pw
    w = [1.0 / (d @ d + eps)]
    vidpos = S.pos[vid].astype(np.float64)
    for vn in S.fan(vid)[:40]:
        nn = S.nn[vn]
        d = S.pinit[nn] - pw
        w.append(1.0 / (d @ d + eps))
        if S.is_feature[nn]:
            off = nn - S.n_free
            tg = S.pos[S.fan(nn)].astype(np.float64) - vidpos
            arg = int(np.argmin(np.einsum("ij,ij->i", tg, tg)))     # the first of equal minima
            sid = int(S.table[off, arg])
            assert sid != -1, (vid, nn, arg)
            N = S.fits[S.frow[off] + sid]
        else:
            N = S.ls[nn]
        q = pw - N[0:3]
        lu, lv = q @ N[3:6], q @ N[6:9]
        pts.append(_l2w(N, lu, lv, _poly(N, lu, lv)))
    w = np.asarray(w)
    return (np.asarray(pts) * (w / w.sum())[:, None]).sum(axis=0)


def feature_neighbours(S: SurfaceState, vid: int):
    """feature neighbours that have exactly two feature neighbours themselves
    (dmo/DMOTetImplCub.h:380-396)"""
    out = []
    for d in S.fan(vid):
        if S.is_feature[d] and sum(1 for d2 in S.fan(d) if S.is_feature[d2]) == 2:
            out.append(int(d))
    return out


def project1d(S: SurfaceState, vid: int, nb, u: float) -> np.ndarray:
    """localSurfaceEstimation1d (dmo/Surface1D.h:54-93): the point (u, a u^2)
    of the NN's curve and its projections onto the curves of the neighbours'
    NNs, blended as in 2D.  In FP64."""
    eps = float(EPS)
    L = S.ls1[S.nn[vid]]
    pw = L[0:3] + u * L[3:6] + (L[9] * u * u) * L[6:9]
    pts = [pw]
    d = S.pinit[S.nn[vid]] - pw
    w = [1.0 / (d @ d + eps)]
    for vn in nb:
        nn = S.nn[vn]
        d = S.pinit[nn] - pw
        w.append(1.0 / (d @ d + eps))
        N = S.ls1[nn]
        lu = (pw - N[0:3]) @ N[3:6]
        pts.append(N[0:3] + lu * N[3:6] + (N[9] * lu * lu) * N[6:9])
    w = np.asarray(w)
    return (np.asarray(pts) * (w / w.sum())[:, None]).sum(axis=0)


def _norm32(d):
    d = np.asarray(d, F)
    return np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])


def _dot32(a, b):
    a, b = np.asarray(a, F), np.asarray(b, F)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def search_surface2d(S: SurfaceState, vid: int, nq=8, depth=3, grid_scale=0.5) -> Result:
    """k_optimizeHierarchical2DTet / 2DHex (dmo/DMOTetImplCub.h:141-322,
    dmo/DMOHexImplCub.h:141-318, USE_SURF_OF_NN = false): (u, v) window
    around the vertex's local coordinates in its NN's frame, half-width
    s_d * the largest (tet) or smallest (hex) distance to the volume ring
    points; thread t holds A = (t/nq, t%nq) and B = (t/nq + 4, t%nq) and keeps
    A unless q_B > q_A; after a move the local position is the chosen (u, v);
    written iff q(old) < the last level's maximum."""
    kind = S.kind
    ring = S.vring(vid)
    x0 = S.pos[vid]
    dist = _norm32(ring - x0)
    dist = dist.min() if kind == "hex" else dist.max()
    L32 = S.ls[S.nn[vid]].astype(F)
    c = x0 - L32[0:3]
    cu, cv = _dot32(L32[3:6], c), _dot32(L32[6:9], c)
    aff = F(1.0) / F(nq - 1)
    T = nq * nq // 2
    t = np.arange(T)
    i1, j1 = t // nq, t % nq
    i2 = i1 + 4
    cur = x0.astype(np.float64)
    ds = F(grid_scale)
    margin = np.inf
    key, qm = 0, None
    ab_tie = False
    for _ in range(depth):
        umax, umin = ds * dist, -ds * dist
        u1 = cu + _grid(aff, i1, nq, umin, umax)
        v1 = cv + _grid(aff, j1, nq, umin, umax)
        u2 = cu + _grid(aff, i2, nq, umin, umax)
        pa = np.array([project2d(S, vid, u1[k], v1[k]) for k in range(T)])
        pb = np.array([project2d(S, vid, u2[k], v1[k]) for k in range(T)])
        qa, qb = ring_quality(kind, ring, pa, S.corner8), ring_quality(kind, ring, pb, S.corner8)
        qlast = float(ring_quality(kind, ring, cur, S.corner8))
        takeb = np.array([gt(b, a) for a, b in zip(qa, qb)])
        qm = np.where(takeb, qb, qa)
        key = _argmax_threads(qm)
        margin = min(margin, _top_gap(np.concatenate([qa, qb])), rel_gap(qm[key], qlast))
        if gt(qm[key], qlast):
            cu, cv = (u2[key], v1[key]) if takeb[key] else (u1[key], v1[key])
            cur = project2d(S, vid, cu, cv)
            ab_tie = ab_tie or eq(qa[key], qb[key])
        ds = ds * (F(2.0) / F(nq - 1))
    qold = float(ring_quality(kind, ring, x0, S.corner8))
    margin = min(margin, rel_gap(qm[key], qold))
    written = gt(qm[key], qold)
    return Result(cur if written else x0.astype(np.float64), written, margin, {"uv": (F(cu), F(cv)), "ab_tie": ab_tie})


def search_feature1d(S: SurfaceState, vid: int, nq=8, depth=3, grid_scale=0.5) -> Result:
    """k_optimizeHierarchical1DTet / 1DHex (dmo/DMOTetImplCub.h:326-494,
    dmo/DMOHexImplCub.h:320-489): nq threads, candidate u = the local u +
    affine grid, half-width s_d * the largest (tet) or smallest (hex) distance
    to the volume ring points; the block maximum per level; written iff
    q(old) < the last level's maximum.  Only vertices eligible for the 1D
    search move."""
    kind = S.kind
    x0 = S.pos[vid]
    if not S.opt1d[vid]:
        return Result(x0.astype(np.float64), False, np.inf, {"uv": None})
    ring = S.vring(vid)
    dist = _norm32(ring - x0)
    dist = dist.min() if kind == "hex" else dist.max()
    nb = feature_neighbours(S, vid)
    L32 = S.ls1[S.nn[vid]].astype(F)
    cu = _dot32(L32[3:6], x0 - L32[0:3])
    aff = F(1.0) / F(nq - 1)
    i1 = np.arange(nq)
    cur = x0.astype(np.float64)
    ds = F(grid_scale)
    margin = np.inf
    key, q = 0, None
    for _ in range(depth):
        umax, umin = ds * dist, -ds * dist
        u = cu + _grid(aff, i1, nq, umin, umax)
        pw = np.array([project1d(S, vid, nb, uu) for uu in u])
        q = ring_quality(kind, ring, pw, S.corner8)
        qlast = float(ring_quality(kind, ring, cur, S.corner8))
        key = _argmax_threads(q)
        margin = min(margin, _top_gap(q), rel_gap(q[key], qlast))
        if gt(q[key], qlast):
            cu = u[key]
            cur = project1d(S, vid, nb, cu)
        ds = ds * (F(2.0) / F(nq - 1))
    qold = float(ring_quality(kind, ring, x0, S.corner8))
    margin = min(margin, rel_gap(q[key], qold))
    written = gt(q[key], qold)
    return Result(cur if written else x0.astype(np.float64), written, margin, {"uv": F(cu)})


def nearest_same_class(pos, pinit, is_feature):
    """For every surface vertex, the nearest initial surface position among the
    vertices of its own class (feature or not), by exhaustive search: what the
    walk of k_updateNearestNeighbor (dmo/DMOCommon.cu:11-43) finds whenever
    it is not stopped by a local minimum.  Returns (index, squared distance
    to it, squared distances to all)."""
    pos = np.asarray(pos, np.float64)
    pinit = np.asarray(pinit, np.float64)
    d2 = ((pos[:, None, :] - pinit[None, :, :]) ** 2).sum(axis=2)
    same = is_feature[:, None] == is_feature[None, :]
    d2 = np.where(same, d2, np.inf)
    idx = np.argmin(d2, axis=1)
    return idx, d2[np.arange(len(pos)), idx], d2


# --------------------------------------------------- learned proposal window
def learned_window(kind: str, x0, ring, c, depth_l, frac, grid_scale=0.5, corner8=False) -> Result:
    """The learned proposal's window search, restated from its specification
    (this build's own search; no reference counterpart): 4 x 4 x 4
    candidates i + 4j + 16k (the lowest index wins ties) on `frac` times the
    plain first-level window (per-axis max extent for tets, min for hexes)
    around the proposal c, `depth_l` levels shrinking by 2/3; the centre
    moves to a level's best candidate when it beats the best quality so far,
    which starts at q(x0).  extra["status"]: 1 moved, 0 stays (a level's best
    ties a positive q(x0), or a hex vertex whose sub-tets without it already
    fail the shape metric's thresholds), -1 falls back to the full search."""
    x0 = np.asarray(x0, F).reshape(3)
    pts = np.asarray(ring, F).reshape(-1, 3)
    d = np.abs(pts - x0)
    ext = d.min(axis=0) if kind == "hex" else d.max(axis=0)
    nl = 4
    aff = F(1.0) / F(nl - 1)
    idx = np.arange(nl ** 3)
    ii, jj, kk = idx % nl, (idx // nl) % nl, idx // (nl * nl)
    q0 = float(ring_quality(kind, pts, x0, corner8))
    cur = np.asarray(c, F).reshape(3).copy()
    qbest = q0
    moved = tie = False
    margin = np.inf
    ds = F(frac) * F(grid_scale)
    for _ in range(depth_l):
        gmax = cur + ds * ext
        gmin = cur - ds * ext
        cand = np.stack([_grid(aff, ii, nl, gmin[0], gmax[0]), _grid(aff, jj, nl, gmin[1], gmax[1]),
                         _grid(aff, kk, nl, gmin[2], gmax[2])], axis=1)
        q = ring_quality(kind, pts, cand, corner8)
        b = 0
        for k in range(1, len(q)):
            if gt(q[k], q[b]):
                b = k
        margin = min(margin, _top_gap(q), rel_gap(q[b], qbest), rel_gap(q[b], q0))
        if eq(q[b], q0) and q0 > 0.0:
            tie = True
        if gt(q[b], qbest):
            cur = cand[b].copy()
            qbest = q[b]
            moved = True
        ds = ds * (F(2.0) / F(nl - 1))
    pinned = False
    if kind == "hex":
        els = pts.reshape(-1, 7, 3).astype(np.float64)
        P = np.concatenate([np.broadcast_to(x0.astype(np.float64), (len(els), 1, 3)), els], axis=1)
        det, jn = hex_subtets(P)
        free = [1, 2, 3, 6]                         # the sub-tets without vertex 0
        pinned = bool(np.any((det[:, free] < float(EPS)) | (jn[:, free] < float(EPS))))
    status = 1 if moved else (0 if (tie or pinned) else -1)
    return Result(cur if moved else x0, moved, margin, {"status": status})


# ------------------------------------------------------ surface-mesh fans
def _fan_points(S: SurfaceState, vid: int, quad: bool):
    """the fan of vid as the kernels load it (dmo/DMOTriImplCub.h:77-90,
    dmo/DMOQuadImplCub.h:82-103): the halfedge targets r_i and, for quads,
    after each the target d_i of the halfedge that follows vid -> r_i in its
    face (a boundary halfedge has no face: its own target stands in, where
    the reference reads out of range); FP32 positions"""
    h0, h1 = S.he_row[vid], S.he_row[vid + 1]
    pts = []
    for h in range(h0, h1):
        pts.append(S.pos[S.he_target[h]])
        if quad:
            nx = S.he_next[h]
            pts.append(S.pos[S.he_target[nx if nx >= 0 else h]])
    return np.asarray(pts, F)


def _q_fan(quad: bool, pts, ne: int, p, n=None):
    """qualityShapeMRMTri / qualityShapeMRMQuad over ne fan elements with p as
    vertex 0 (dmo/DMOQualityTri.h:313-343, dmo/DMOQualityQuad.h:55-90), in
    FP64; with a normal n the signed variants meanRatioMetricTri(v, n)
    (DMOQualityTri.h:280-296: 2 sqrt(3) |N| sign(N/|N| . n) / sum l^2) and
    meanRatioMetricQuad(v, n) (DMOQualityQuad.h:31-50: -1 when a corner
    triangle's unit normal has . n < 0.7)"""
    P = np.asarray(pts, np.float64)
    p = np.asarray(p, np.float64)
    q = np.inf
    for k in range(ne):
        if quad:
            v = [p, P[2 * k], P[2 * k + 1], P[2 * k + 2]]
            e = [v[1] - v[0], v[2] - v[1], v[3] - v[2], v[0] - v[3]]
            l = sum(x @ x for x in e)
            n1, n2 = np.cross(e[1], -e[0]), np.cross(e[3], -e[2])
            a1, a2 = np.linalg.norm(n1), np.linalg.norm(n2)
            if n is not None and ((n1 / a1) @ n < 0.7 or (n2 / a2) @ n < 0.7):
                val = -1.0
            else:
                val = 2.0 * (a1 + a2) / l
        else:
            v = [p, P[k], P[k + 1]]
            e0, e1, e2 = v[1] - v[0], v[2] - v[0], v[2] - v[1]
            l = e0 @ e0 + e1 @ e1 + e2 @ e2
            N = np.cross(e0, e1)
            a = np.linalg.norm(N)
            sgn = 1.0 if n is None else float(np.sign((N / a) @ n)) if a > 0 else 0.0
            val = 2.0 * np.sqrt(3.0) * a * sgn / l
        q = min(q, val)
    return q


def _fan_setup(S: SurfaceState, vid: int, open_fan: bool):
    quad = S.kind == "quad"
    pts = _fan_points(S, vid, quad)
    nh = S.he_row[vid + 1] - S.he_row[vid]
    ne = nh - 1 if open_fan else nh
    if not open_fan:
        pts = np.concatenate([pts, pts[:1]])                   # r_n = r_0
    dist = _norm32(pts - S.pos[vid]).max()
    n = S.vnorm[S.nn[vid]].astype(np.float64)
    return quad, pts, ne, dist, n


def search_fan2d(S: SurfaceState, vid: int, nq=8, depth=3, grid_scale=0.5) -> Result:
    """k_optimizeHierarchical2D of triangle / quad surface meshes
    (dmo/DMOTriImplCub.h:25-200, dmo/DMOQuadImplCub.h:23-212): as
    search_surface2d, with the closed fan's metric signed by the NN's vertex
    normal as the score, the largest distance to the fan points as the
    window, and the unsigned metric of the start position in the write check."""
    quad, pts, ne, dist, nrm = _fan_setup(S, vid, False)
    x0 = S.pos[vid]
    L32 = S.ls[S.nn[vid]].astype(F)
    c = x0 - L32[0:3]
    cu, cv = _dot32(L32[3:6], c), _dot32(L32[6:9], c)
    aff = F(1.0) / F(nq - 1)
    T = nq * nq // 2
    t = np.arange(T)
    i1, j1 = t // nq, t % nq
    i2 = i1 + 4
    cur = x0.astype(np.float64)
    ds = F(grid_scale)
    margin = np.inf
    key, qm = 0, None
    for _ in range(depth):
        umax, umin = ds * dist, -ds * dist
        u1 = cu + _grid(aff, i1, nq, umin, umax)
        v1 = cv + _grid(aff, j1, nq, umin, umax)
        u2 = cu + _grid(aff, i2, nq, umin, umax)
        qa = np.array([_q_fan(quad, pts, ne, project2d(S, vid, u1[k], v1[k]), nrm) for k in range(T)])
        qb = np.array([_q_fan(quad, pts, ne, project2d(S, vid, u2[k], v1[k]), nrm) for k in range(T)])
        qlast = _q_fan(quad, pts, ne, cur, nrm)
        takeb = np.array([gt(b, a) for a, b in zip(qa, qb)])
        qm = np.where(takeb, qb, qa)
        key = _argmax_threads(qm)
        margin = min(margin, _top_gap(np.concatenate([qa, qb])), rel_gap(qm[key], qlast))
        if gt(qm[key], qlast):
            cu, cv = (u2[key], v1[key]) if takeb[key] else (u1[key], v1[key])
            cur = project2d(S, vid, cu, cv)
        ds = ds * (F(2.0) / F(nq - 1))
    qold = _q_fan(quad, pts, ne, x0)
    margin = min(margin, rel_gap(qm[key], qold))
    written = gt(qm[key], qold)
    return Result(cur if written else x0.astype(np.float64), written, margin, {"uv": (F(cu), F(cv))})


def search_fan1d(S: SurfaceState, vid: int, nq=8, depth=3, grid_scale=0.5) -> Result:
    """k_optimizeHierarchical1D of triangle / quad surface meshes
    (dmo/DMOTriImplCub.h:205-378, dmo/DMOQuadImplCub.h:216-402): the feature
    curve search of search_feature1d scored on the fan (open on a surface
    boundary) signed by the NN's vertex normal; write check unsigned."""
    x0 = S.pos[vid]
    if not S.opt1d[vid]:
        return Result(x0.astype(np.float64), False, np.inf, {"uv": None})
    quad, pts, ne, dist, nrm = _fan_setup(S, vid, bool(S.is_b1d[vid]))
    nb = feature_neighbours(S, vid)
    L32 = S.ls1[S.nn[vid]].astype(F)
    cu = _dot32(L32[3:6], x0 - L32[0:3])
    aff = F(1.0) / F(nq - 1)
    i1 = np.arange(nq)
    cur = x0.astype(np.float64)
    ds = F(grid_scale)
    margin = np.inf
    key, q = 0, None
    for _ in range(depth):
        umax, umin = ds * dist, -ds * dist
        u = cu + _grid(aff, i1, nq, umin, umax)
        q = np.array([_q_fan(quad, pts, ne, project1d(S, vid, nb, uu), nrm) for uu in u])
        qlast = _q_fan(quad, pts, ne, cur, nrm)
        key = _argmax_threads(q)
        margin = min(margin, _top_gap(q), rel_gap(q[key], qlast))
        if gt(q[key], qlast):
            cu = u[key]
            cur = project1d(S, vid, nb, cu)
        ds = ds * (F(2.0) / F(nq - 1))
    qold = _q_fan(quad, pts, ne, x0)
    margin = min(margin, rel_gap(q[key], qold))
    written = gt(q[key], qold)
    return Result(cur if written else x0.astype(np.float64), written, margin, {"uv": F(cu)})


def search_flat1d(S: SurfaceState, vid: int, nq=8, depth=3, grid_scale=0.5) -> Result:
    """k_optimizeHierarchical1DFlat (dmo/DMOTriFlatImplCub.h:158-333): a
    boundary vertex of a flat triangle mesh follows its boundary curve; the
    score is the 2D mean ratio over its fan (closed unless the vertex lies on
    a surface boundary), the window the largest distance to the fan targets;
    written iff q(old) < the last level's maximum."""
    x0 = S.pos[vid]
    if not S.opt1d[vid]:
        return Result(x0.astype(np.float64), False, np.inf, {"uv": None})
    r = S.pos[S.fan(vid)].astype(F)
    dist = _norm32(r - x0).max()
    ring = r if S.is_b1d[vid] else np.concatenate([r, r[:1]])
    els = np.stack([ring[:-1], ring[1:]], axis=1).astype(np.float64)     # (r_k, r_k+1)

    def score(p):
        P = np.concatenate([np.broadcast_to(np.asarray(p, np.float64), (len(els), 1, 3)), els], axis=1)
        return float(q_tri2d(P).min())

    nb = feature_neighbours(S, vid)
    L32 = S.ls1[S.nn[vid]].astype(F)
    cu = _dot32(L32[3:6], x0 - L32[0:3])
    aff = F(1.0) / F(nq - 1)
    i1 = np.arange(nq)
    cur = x0.astype(np.float64)
    ds = F(grid_scale)
    margin = np.inf
    key, q = 0, None
    for _ in range(depth):
        umax, umin = ds * dist, -ds * dist
        u = cu + _grid(aff, i1, nq, umin, umax)
        q = np.array([score(project1d(S, vid, nb, uu)) for uu in u])
        qlast = score(cur)
        key = _argmax_threads(q)
        margin = min(margin, _top_gap(q), rel_gap(q[key], qlast))
        if gt(q[key], qlast):
            cu = u[key]
            cur = project1d(S, vid, nb, cu)
        ds = ds * (F(2.0) / F(nq - 1))
    qold = score(x0)
    margin = min(margin, rel_gap(q[key], qold))
    written = gt(q[key], qold)
    return Result(cur if written else x0.astype(np.float64), written, margin, {"uv": F(cu)})
