"""GPU tests of the learned proposal: the BF16 MFMA MLP against the FP64
restatement (forward and gradients), training (loss decreases, seeds
reproduce), the recorded training pairs against the feature encoding, the
fallback rule, and what the trained network gains over the same windowed
search without it.

Tolerances: forward max |dY| <= 2e-3 + 5e-3 |Y| (BF16 activations rounded on
both sides; an FP32-vs-FP64 accumulation difference can flip one rounding);
gradients per parameter block ||g - g_ref|| <= 3e-2 ||g_ref|| (the backward
pass rounds dY, dZ2, dZ1 to BF16 for the MFMA products).
"""
import numpy as np
import pytest

import mlp_ref
from volumesurfacediscretemeshoptimization_amd import (DMODataset, DMOHexClass, DMOMeshHexFactory, DMOMeshTetFactory,
                                                       DMOMLP, DMOTetClass, VsdmoError)
from volumesurfacediscretemeshoptimization_amd import synth

pytestmark = pytest.mark.gpu


def _xy(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(n, 96)) * 0.6).astype(np.float32)
    T = (rng.normal(size=(n, 3)) * 0.1).astype(np.float32)
    return X, T


def test_mlp_forward_matches_reference():
    mlp = DMOMLP(seed=3)
    p = mlp.params().astype(np.float64)
    X, _ = _xy(333, 1)
    Y = mlp.forward(X).astype(np.float64)
    Yr, _ = mlp_ref.forward(p, X, emulate=True)
    err = np.abs(Y - Yr)
    assert np.all(err <= 2e-3 + 5e-3 * np.abs(Yr)), err.max()
    # without emulation the BF16 network is still close to the FP64 one
    Y64, _ = mlp_ref.forward(p, X, emulate=False)
    assert np.abs(Y - Y64).max() < 0.05
    mlp.close()


def test_mlp_gradients_match_fp64():
    mlp = DMOMLP(seed=4)
    p = mlp.params().astype(np.float64)
    X, T = _xy(257, 2)
    loss, g = mlp.gradients(X, T)
    lr_, gr = mlp_ref.loss_and_grad(p, X, T, emulate=True)
    assert abs(loss - lr_) <= 1e-3 * abs(lr_), (loss, lr_)
    for name, sl in mlp_ref.block_slices().items():
        a, b = g[sl].astype(np.float64), gr[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < 3e-2, (name, rel)
    mlp.close()


def _record(kind, sizes_seeds, iters, cap):
    ds = DMODataset(cap)
    for n, seed in sizes_seeds:
        if kind == "tet":
            xyz, cells = synth.tet_cube(n, seed=seed)
            m = DMOMeshTetFactory.create(xyz, cells)
            o = DMOTetClass(m, n_iter=iters, fixed_boundary=True)
        else:
            xyz, cells = synth.hex_cube(n, seed=seed, jitter=0.2)
            m = DMOMeshHexFactory.create(xyz, cells)
            o = DMOHexClass(m, n_iter=iters, fixed_boundary=True)
        o.record(ds, iters)
        o.close()
        m.close()
    return ds


def test_dataset_pairs_bit_exact(oracle):
    xyz, tets = synth.tet_cube(8, seed=5)
    m = DMOMeshTetFactory.create(xyz, tets)
    o = DMOTetClass(m, n_iter=1, fixed_boundary=True)
    p0 = m.positions(original_order=False)
    ds = DMODataset(100000)
    o.record(ds, 1)
    p1 = m.positions(original_order=False)
    c = m.counts()
    assert len(ds) == c["n_inner"]
    # recording does not change the trajectory
    m2 = DMOMeshTetFactory.create(xyz, tets)
    o2 = DMOTetClass(m2, n_iter=1, fixed_boundary=True)
    o2.iterate(1)
    assert np.array_equal(m2.positions(original_order=False), p1)
    # first colour class: features from the start positions, targets from the moves
    off = m.color_offsets("inner")
    lo, hi = int(off[0]), int(off[1])
    X, T = ds.get(0, hi - lo)
    om = oracle.OracleMesh("tet", xyz, tets)
    row, val = om.ring()
    assert np.array_equal(om.positions(), p0)
    for i, v in enumerate(range(lo, hi)):
        nb = mlp_ref.canonical_neighbours(row, val, 3, v)
        f, L = mlp_ref.features(p0, nb, v)
        assert np.array_equal(X[i], f), v
        t = ((p1[v] - p0[v]) / L).astype(np.float32)
        assert np.array_equal(T[i], t), v


def test_training_reduces_loss_and_is_deterministic():
    ds = _record("tet", [(10, 100), (10, 101)], 4, 1 << 20)
    n = len(ds)
    assert n == 2 * 4 * 9 ** 3
    a, b, c = DMOMLP(seed=7), DMOMLP(seed=7), DMOMLP(seed=8)
    la = a.train(ds, epochs=6, batch=256, lr=1e-3, seed=11)
    lb = b.train(ds, epochs=6, batch=256, lr=1e-3, seed=11)
    c.train(ds, epochs=1, batch=256, lr=1e-3, seed=11)
    assert la[-1] < 0.7 * la[0], la
    assert np.array_equal(la, lb)
    assert np.array_equal(a.params(), b.params())
    assert not np.array_equal(a.params(), c.params())


def test_fallback_rule_bad_proposal():
    """A network whose proposal lies far outside every one-ring: no window
    candidate beats q(v), so every vertex falls back to the plain search and
    the positions equal the plain run bit for bit."""
    xyz, tets = synth.tet_cube(10, seed=12, jitter=0.2)
    mlp = DMOMLP(seed=1)
    p = np.zeros(mlp.n_params, dtype=np.float32)
    p[-3:] = 50.0                     # b3: displacement 50 L in every axis
    mlp.set_params(p)
    m1 = DMOMeshTetFactory.create(xyz, tets)
    m2 = DMOMeshTetFactory.create(xyz, tets)
    o1 = DMOTetClass(m1, n_iter=2, fixed_boundary=True)
    o2 = DMOTetClass(m2, n_iter=2, fixed_boundary=True)
    assert o1.getQuality() > 0
    o2.set_proposal(mlp, mode=1, depth=2, window_frac=0.25)
    o1.iterate(2)
    o2.iterate(2)
    assert np.array_equal(m1.positions(original_order=False), m2.positions(original_order=False))
    st = o2.stats()
    assert st["searches_learned"] == 2 * m1.counts()["n_inner"]
    assert st["fallbacks"] == st["searches_learned"]


def test_identity_proposal_monotone():
    xyz, tets = synth.tet_cube(12, seed=13)
    m = DMOMeshTetFactory.create(xyz, tets)
    o = DMOTetClass(m, n_iter=5, fixed_boundary=True)
    o.set_proposal(None, mode=2, depth=2, window_frac=0.25)
    q = [o.getQuality()]
    for _ in range(5):
        o.doIteration()
        q.append(o.getQuality())
    assert all(b >= a for a, b in zip(q, q[1:])), q
    st = o.stats()
    assert 0 < st["fallbacks"] < st["searches_learned"]


_ABLATION = [("tet", 4, 0.25, "immediate", 1), ("hex", 4, 0.25, "immediate", 1),
             # the shipped windows and policy (bench default), with the network
             # alone (mode 1) and with the vertex's own window as a second chance (mode 3)
             ("tet", 1, 0.025, "deferred", 1), ("hex", 1, 2.0, "deferred", 1),
             ("tet", 1, 0.025, "deferred", 3), ("hex", 1, 2.0, "deferred", 3)]


@pytest.fixture(scope="module")
def ablation_nets():
    out = {}
    for kind in ("tet", "hex"):
        n = 14 if kind == "tet" else 12
        ds = _record(kind, [(n, 200 + s) for s in range(8)], 8, 1 << 21)
        mlp = DMOMLP(seed=5)
        loss = mlp.train(ds, epochs=20, batch=1024, lr=1e-3, seed=6)
        assert loss[-1] < loss[0]
        ds.close()
        out[kind] = mlp
    yield out
    for m in out.values():
        m.close()


@pytest.mark.parametrize("kind,depth,frac,policy,mode", _ABLATION)
def test_learned_reaches_plain_quality_and_ablation(ablation_nets, kind, depth, frac, policy, mode):
    """Trained on seeded jittered meshes, the proposal reaches the plain
    run's final minimum quality within 1e-3, and gains over the same windowed
    search around the vertex itself (mode 2) at the same window and policy:
    fewer iterations to that quality, or as many with fewer fallbacks to the
    full search."""
    n = 14 if kind == "tet" else 12
    mlp = ablation_nets[kind]
    if kind == "tet":
        xyz, cells = synth.tet_cube(n, seed=999)
        fac, cls = DMOMeshTetFactory, DMOTetClass
    else:
        xyz, cells = synth.hex_cube(n, seed=999, jitter=0.2)
        fac, cls = DMOMeshHexFactory, DMOHexClass
    iters = 10
    mp = fac.create(xyz, cells)
    op = cls(mp, n_iter=iters, fixed_boundary=True)
    op.iterate(iters)
    q_plain = op.getQuality()
    res = {}
    for md in (mode, 2):
        m = fac.create(xyz, cells)
        o = cls(m, n_iter=3 * iters, fixed_boundary=True)
        o.set_proposal(mlp if md != 2 else None, mode=md, depth=depth, window_frac=frac, fallback=policy)
        q = [o.getQuality()]
        it = 0
        while it < 2 * iters and q[-1] < q_plain - 1e-3:
            o.doIteration()
            it += 1
            q.append(o.getQuality())
        st = o.stats()
        res[md] = (it, q[-1], st["fallbacks"] / max(1, st["searches_learned"]))
        assert all(b >= a for a, b in zip(q, q[1:]))
        o.close()
        m.close()
    print(kind, depth, frac, policy, "plain", q_plain, f"mode {mode}", res[mode], "identity", res[2])
    assert res[mode][1] >= q_plain - 1e-3, res
    assert res[mode][0] < res[2][0] or (res[mode][0] == res[2][0] and res[mode][2] < res[2][2]), res


@pytest.mark.parametrize("fallback", ["immediate", "deferred"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_fused_network_equals_separate_launch(kind, fallback, monkeypatch):
    """The window kernel's own run of the network (16-vertex tiles, partial
    last tile per class) gives the positions, fallback count and quality of
    the separate inference launch bit for bit."""
    mlp = DMOMLP(seed=21)
    p = mlp.params()
    p[-3:] = [0.05, -0.03, 0.02]         # b3: a nonzero mean displacement
    mlp.set_params(p)
    if kind == "tet":
        xyz, cells = synth.tet_cube(13, seed=31)
        fac, cls = DMOMeshTetFactory, DMOTetClass
    else:
        xyz, cells = synth.hex_cube(11, seed=31, jitter=0.2)
        fac, cls = DMOMeshHexFactory, DMOHexClass
    out = []
    for sep in ("1", "0"):
        monkeypatch.setenv("VSDMO_LEARN_SEPARATE_INFERENCE", sep)
        m = fac.create(xyz, cells)
        o = cls(m, n_iter=3, fixed_boundary=True)
        o.set_proposal(mlp, mode=1, fallback=fallback)
        o.iterate(3)
        st = o.stats()
        out.append((m.positions(original_order=False), st["fallbacks"], o.getQuality()))
        o.close()
        m.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and 0 < out[0][1]
    assert out[0][2] == out[1][2]
    mlp.close()


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_deferred_fallback_bad_proposal(kind):
    """Deferred fallback with a proposal far outside every one-ring: one pass
    defers every vertex (none moves), the class's next pass runs their full
    searches and skips their windows, so 2k iterations equal k plain ones
    bit for bit (hexes: pinned vertices are neither deferred nor moved by
    the full search)."""
    if kind == "tet":
        xyz, cells = synth.tet_cube(10, seed=12, jitter=0.2)
        fac, cls = DMOMeshTetFactory, DMOTetClass
    else:
        xyz, cells = synth.hex_cube(9, seed=12, jitter=0.3)
        fac, cls = DMOMeshHexFactory, DMOHexClass
    mlp = DMOMLP(seed=1)
    p = np.zeros(mlp.n_params, dtype=np.float32)
    p[-3:] = 50.0
    mlp.set_params(p)
    m1, m2 = fac.create(xyz, cells), fac.create(xyz, cells)
    o1 = cls(m1, n_iter=2, fixed_boundary=True)
    o2 = cls(m2, n_iter=4, fixed_boundary=True)
    o2.set_proposal(mlp, mode=1, depth=2, window_frac=0.25, fallback="deferred")
    p0 = m2.positions(original_order=False)
    for k in range(2):
        o2.doIteration()
        if k == 0:
            assert np.array_equal(m2.positions(original_order=False), p0)
        o2.doIteration()
        o1.doIteration()
        assert np.array_equal(m1.positions(original_order=False), m2.positions(original_order=False)), k
        assert o1.getQuality() == o2.getQuality()
    st = o2.stats()
    if kind == "tet":
        assert st["fallbacks"] == 2 * m1.counts()["n_inner"]
    assert st["searches_learned"] == 4 * m1.counts()["n_inner"]
    for o, m in ((o1, m1), (o2, m2)):
        o.close()
        m.close()
    mlp.close()


def test_deferred_fallback_monotone_and_policy_checks():
    """Identity proposal with deferred fallback: the minimum quality never
    decreases; the policy needs nq 4 or 8."""
    xyz, tets = synth.tet_cube(12, seed=13)
    m = DMOMeshTetFactory.create(xyz, tets)
    o = DMOTetClass(m, n_iter=6, fixed_boundary=True)
    o.set_proposal(None, mode=2, depth=2, window_frac=0.25, fallback="deferred")
    q = [o.getQuality()]
    for _ in range(6):
        o.doIteration()
        q.append(o.getQuality())
    assert all(b >= a for a, b in zip(q, q[1:])), q
    st = o.stats()
    assert 0 < st["fallbacks"] < st["searches_learned"]
    with pytest.raises(ValueError):
        o.set_fallback("later")
    o.close()
    o6 = DMOTetClass(m, n_iter=1, fixed_boundary=True, nq=6)
    with pytest.raises(VsdmoError):
        o6.set_fallback("deferred")
    o6.close()
    # nq 4: the deferred full searches run the 4^3 grid (one candidate per lane)
    o4 = DMOTetClass(m, n_iter=4, fixed_boundary=True, nq=4)
    o4.set_proposal(None, mode=2, depth=1, window_frac=0.25, fallback="deferred")
    q = [o4.getQuality()]
    for _ in range(4):
        o4.doIteration()
        q.append(o4.getQuality())
    assert all(b >= a for a, b in zip(q, q[1:])), q
    assert o4.stats()["fallbacks"] > 0
    o4.close()
    m.close()


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_default_window_deferred_reaches_plain_quality(kind):
    """The defaults the bench measures (window per element type, deferred
    fallback): trained on seeded meshes, the learned run reaches the plain
    run's final minimum and mean quality (within 1e-3) in fewer iterations
    than the plain run's 10, with a non-decreasing minimum."""
    n = 14 if kind == "tet" else 12
    ds = _record(kind, [(n, 300 + s) for s in range(8)], 8, 1 << 21)
    mlp = DMOMLP(seed=5)
    mlp.train(ds, epochs=20, batch=1024, lr=1e-3, seed=6)
    if kind == "tet":
        xyz, cells = synth.tet_cube(n, seed=998)
        fac, cls = DMOMeshTetFactory, DMOTetClass
    else:
        xyz, cells = synth.hex_cube(n, seed=998, jitter=0.2)
        fac, cls = DMOMeshHexFactory, DMOHexClass
    mp = fac.create(xyz, cells)
    op = cls(mp, n_iter=10, fixed_boundary=True)
    op.iterate(10)
    qp, mq = op.getQuality(), op.getMeanQuality()
    m = fac.create(xyz, cells)
    o = cls(m, n_iter=20, fixed_boundary=True)
    o.set_proposal(mlp, mode=1, fallback="deferred")
    q = [o.getQuality()]
    it = 0
    while it < 10 and not (q[-1] >= qp - 1e-3 and o.getMeanQuality() >= mq - 1e-3):
        o.doIteration()
        it += 1
        q.append(o.getQuality())
    print(kind, "plain", qp, mq, "learned", it, q[-1], o.getMeanQuality(), o.stats()["fallbacks"])
    assert q[-1] >= qp - 1e-3 and o.getMeanQuality() >= mq - 1e-3, (it, q[-1], qp)
    assert it < 10
    assert all(b >= a for a, b in zip(q, q[1:])), q
    for x in (o, m, op, mp, mlp):
        x.close()


def test_deferred_with_surface_passes():
    """A valid hex mesh with the surface free: the learned inner passes
    (deferred fallback) are followed by the surface and feature passes every
    iteration; the minimum quality never decreases and the surface searches
    run."""
    xyz, hexes = synth.hex_cube(10, seed=21, jitter=0.15)
    m = DMOMeshHexFactory.create(xyz, hexes)
    o = DMOHexClass(m, n_iter=4, fixed_boundary=False)
    assert o.getQuality() > 0
    o.set_proposal(None, mode=2, fallback="deferred")
    q = [o.getQuality()]
    for _ in range(4):
        o.doIteration()
        q.append(o.getQuality())
    assert all(b >= a for a, b in zip(q, q[1:])), q
    st = o.stats()
    assert st["searches_surface"] > 0 and st["searches_learned"] == 4 * m.counts()["n_inner"]
    o.close()
    m.close()


def test_deferred_lists_dropped_by_recording():
    """Deferred learned iterations, a recorded (plain) iteration, then deferred
    iterations again: the recording drops the pending deferrals, so no vertex
    is searched twice at once -- the run is reproducible bit for bit and its
    minimum quality never decreases."""
    xyz, tets = synth.tet_cube(12, seed=17)
    mlp = DMOMLP(seed=3)
    p = mlp.params()
    p[-3:] = [0.04, -0.02, 0.03]
    mlp.set_params(p)
    runs = []
    for _ in range(2):
        m = DMOMeshTetFactory.create(xyz, tets)
        o = DMOTetClass(m, n_iter=10, fixed_boundary=True)
        o.set_proposal(mlp, mode=1, fallback="deferred")
        q = [o.getQuality()]
        for k in range(5):
            if k == 2:
                ds = DMODataset(100000)
                o.record(ds, 1)
                ds.close()
            else:
                o.doIteration()
            q.append(o.getQuality())
        assert all(b >= a for a, b in zip(q, q[1:])), q
        assert o.stats()["fallbacks"] > 0
        runs.append(m.positions(original_order=False))
        o.close()
        m.close()
    assert np.array_equal(runs[0], runs[1])
    mlp.close()


def test_fallback_policy_switch_mid_run():
    """Deferred -> immediate -> deferred on one optimiser: every switch starts
    from clean lists (the immediate policy's list lengths are zeroed, pending
    deferrals are dropped); the minimum quality never decreases and the run
    is reproducible bit for bit."""
    xyz, tets = synth.tet_cube(12, seed=19)
    mlp = DMOMLP(seed=4)
    p = mlp.params()
    p[-3:] = [0.03, 0.02, -0.04]
    mlp.set_params(p)
    runs = []
    for _ in range(2):
        m = DMOMeshTetFactory.create(xyz, tets)
        o = DMOTetClass(m, n_iter=12, fixed_boundary=True)
        o.set_proposal(mlp, mode=1, fallback="deferred")
        q = [o.getQuality()]
        for k, pol in enumerate(["deferred", "deferred", "immediate", "immediate", "deferred", "deferred"]):
            if k in (2, 4):
                o.set_fallback(pol)
            o.doIteration()
            q.append(o.getQuality())
        assert all(b >= a for a, b in zip(q, q[1:])), q
        st = o.stats()
        assert 0 < st["fallbacks"] < st["searches_learned"]
        runs.append(m.positions(original_order=False))
        o.close()
        m.close()
    assert np.array_equal(runs[0], runs[1])
    mlp.close()
