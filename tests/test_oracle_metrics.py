"""Pins the oracle's canonical FP32 metrics (CPU only).

Checks against closed forms, an FP64 restatement of the reference formulas
and invariances (rotation, translation, power-of-two scaling) that a dropped
term, a wrong sign or a transposed operand would break.
"""
import numpy as np
import pytest

S3 = np.sqrt(3.0)


def regular_tet():
    return np.array([[0, 0, 0], [1, 0, 0], [0.5, S3 / 2, 0], [0.5, S3 / 6, np.sqrt(2 / 3)]], np.float32)


def test_cbrt_accuracy(oracle):
    rng = np.random.default_rng(1)
    xs = (rng.random(20000) * 7 + 1) * 2.0 ** rng.integers(-120, 120, 20000)
    xs = xs.astype(np.float32)
    worst = 0.0
    for x in xs[:4000]:
        y = oracle.cbrt(float(x))
        t = np.cbrt(np.float64(x))
        worst = max(worst, abs(y - t) / np.spacing(np.float32(t)))
    assert worst < 1.0, worst
    assert oracle.cbrt(27.0) == 3.0 and oracle.cbrt(8.0) == 2.0 and oracle.cbrt(0.0) == 0.0
    # denormal input path
    d = float(np.float32(1e-40))
    assert abs(oracle.cbrt(d) / np.cbrt(d) - 1) < 2e-7


def test_tet_closed_forms(oracle):
    assert oracle.q_elem("tet", regular_tet()) == pytest.approx(1.0, abs=2e-7)
    kuhn = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], np.float32)
    # 12*cbrt(9)*cbrt(1/36)/10 = 0.7559526...: the unjittered Kuhn-cube quality
    assert oracle.q_elem("tet", kuhn) == pytest.approx(0.7559526299369237, rel=2e-7)
    # negative orientation returns the signed volume (-sqrt(2)/12 for the regular tet)
    inv = regular_tet()[[0, 2, 1, 3]]
    assert oracle.q_elem("tet", inv) == pytest.approx(-np.sqrt(2) / 12, rel=1e-6)
    # flat tet -> 0
    flat = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    assert oracle.q_elem("tet", flat) == 0.0


def test_hex_closed_forms(oracle):
    cube = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    assert oracle.q_elem("hex", cube) == pytest.approx(1.0, abs=2e-7)
    # bottom-first (raw MEDIT) order is inverted for the metric -> 0
    assert oracle.q_elem("hex", cube[[4, 5, 6, 7, 0, 1, 2, 3]]) == 0.0
    # a box a x b x c: sub-tet dets a*b*c, |J|^2 = a^2+b^2+c^2; principal tet 64abc / 16(a^2+b^2+c^2)
    a, b, c = 1.0, 2.0, 0.5
    box = cube * np.array([a, b, c], np.float32)
    q = oracle.q_elem("hex", box)
    exp = 3 * np.cbrt((a * b * c) ** 2) / (a * a + b * b + c * c)
    assert q == pytest.approx(exp, rel=1e-6)
    assert oracle.q_elem("hex", box, fp64=True) == pytest.approx(exp, rel=1e-12)


def test_tri_closed_forms(oracle):
    eq = np.array([[0, 0, 0], [1, 0, 0], [0.5, S3 / 2, 0]], np.float32)
    assert oracle.q_elem("tri", eq) == pytest.approx(1.0, abs=2e-7)
    right = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    assert oracle.q_elem("tri", right) == pytest.approx(S3 / 2, rel=2e-7)
    # clockwise: returns the signed cross value (-2*area)
    assert oracle.q_elem("tri", right[[0, 2, 1]]) == pytest.approx(-1.0, rel=1e-7)


def _rand_tet(rng):
    while True:
        p = rng.random((4, 3)) * 2 - 1
        if np.linalg.det(p[1:] - p[0]) > 0.05:
            return p.astype(np.float32)


def _rand_hex(rng):
    base = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    return (base + rng.uniform(-0.2, 0.2, base.shape)).astype(np.float32)


@pytest.mark.parametrize("kind", ["tet", "hex", "tri"])
def test_fp32_matches_fp64(oracle, kind):
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(3000):
        if kind == "tet":
            p = _rand_tet(rng)
        elif kind == "hex":
            p = _rand_hex(rng)
        else:
            p = np.zeros((3, 3), np.float32)
            p[:, :2] = rng.random((3, 2))
        q32 = oracle.q_elem(kind, p)
        q64 = oracle.q_elem(kind, p, fp64=True)
        err = abs(q32 - q64)
        # FP32 rounding of the edge vectors (coordinates of magnitude ~1): an
        # absolute 1e-6 floor for slivers plus 5e-6 relative
        assert err < 1e-6 + 5e-6 * abs(q64), (p, q32, q64)
        if q64 > 0.1:
            worst = max(worst, err / q64)
    assert worst < 2e-6, worst


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_invariances(oracle, kind):
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = _rand_tet(rng) if kind == "tet" else _rand_hex(rng)
        q = oracle.q_elem(kind, p)
        # power-of-two scaling changes no rounding: bit-identical
        assert oracle.q_elem(kind, (p * np.float32(8.0)).astype(np.float32)) == q
        # rotation (FP32 rounding of the rotated coordinates -> tolerance)
        th = rng.random() * 6.28
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        q2 = oracle.q_elem(kind, (p.astype(np.float64) @ R.T).astype(np.float32))
        assert q2 == pytest.approx(q, rel=5e-5, abs=1e-6)


def test_ring_score_is_ring_minimum(oracle):
    """Ring score = min over the ring's elements of the element quality."""
    rng = np.random.default_rng(11)
    for _ in range(100):
        x0 = rng.random(3).astype(np.float32)
        ring = (x0 + rng.uniform(-1, 1, (24, 3, 3))).astype(np.float32)
        # orient each element positively around x0
        for e in range(24):
            if np.linalg.det(ring[e] - x0) < 0:
                ring[e] = ring[e][[0, 2, 1]]
        p = (x0 + rng.uniform(-0.1, 0.1, 3)).astype(np.float32)
        s = oracle.ring_score("tet", ring.ravel(), x0, 1.0, p)
        ref = min(oracle.q_elem("tet", np.vstack([p, ring[e]]), fp64=True) for e in range(24))
        assert abs(s - ref) < 1e-6 + 5e-6 * abs(ref), (s, ref)
