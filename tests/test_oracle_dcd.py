"""Tri-tri DCD oracle pinned by the reference's known answers
(tests/test_udcf_dcd.py:191-236 respond full push / half split, :239-286 emit
rules, :289-332 penetration resolution, :335-392 kinetic friction) and by
brute-force / symmetry properties."""
import numpy as np

import oracle

X6 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0],
               [0.0, 0.005, 0.0], [1.0, 0.005, 0.0], [0.0, 0.005, 1.0]], np.float32)
B3 = (1 / 3, 1 / 3, 1 / 3)


def test_respond_solid_side_full_push():
    c = oracle.make_contact([0, 1, 2], [3, 4, 5], 0.01, (0, 1, 0), B3, B3)
    out = oracle.dcd_respond_apply(c, 1, X6, X6, np.array([1, 1, 1, 0, 0, 0]), 1.01, 0.5, 0.4)
    assert np.allclose(out[0:3, 1], 0.01 * 1.01, atol=1e-6)
    assert np.allclose(out[3:6], X6[3:6], atol=1e-7)


def test_respond_both_movable_half_split():
    c = oracle.make_contact([0, 1, 2], [3, 4, 5], 0.01, (0, 1, 0), B3, B3)
    out = oracle.dcd_respond_apply(c, 1, X6, X6, np.ones(6), 1.01, 0.5, 0.4)
    half = 0.01 * 1.01 * 0.5
    assert np.allclose(out[0:3, 1] - X6[0:3, 1], half, atol=1e-6)
    assert np.allclose(out[3:6, 1] - X6[3:6, 1], -half, atol=1e-6)


def test_respond_kinetic_friction():
    prev = X6.copy()
    prev[0:3, 0] -= 0.02
    c = oracle.make_contact([0, 1, 2], [3, 4, 5], 0.01, (0, 1, 0), B3, B3)
    out = oracle.dcd_respond_apply(c, 1, X6, prev, np.array([1, 1, 1, 0, 0, 0]), 1.01, 0.5, 0.4)
    assert np.allclose(out[0:3] - X6[0:3], [-0.002, 0.0101, 0.0], atol=1e-6)
    assert np.allclose(out[3:6], X6[3:6], atol=1e-7)


def test_emit_rules_four_tris():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0.1, 0.01, 0.1], [1, 0.01, 1],
                  [0, 0.02, 0], [1, 0.02, 0], [0, 0.02, 1], [0, 0.03, 0], [1, 0.03, 0], [0, 0.03, 1]],
                 np.float32)
    tris = np.array([[0, 1, 2], [1, 3, 4], [5, 6, 7], [8, 9, 10]], np.int32)
    group = np.array([0, 0, 1, 1], np.int32)
    for allow in (0, 1):
        p = oracle.dcd_pairs(tris, x, group=group, allow_self=allow, thickness=0.1)
        assert set(map(tuple, p)) == {(0, 2), (0, 3), (1, 2), (1, 3)}


def test_penetration_resolves():
    """tilted movable tri through a static one: 8 Jacobi passes halve the depth."""
    mov = np.array([[-0.5, -0.02, -0.3], [0.5, 0.04, -0.3], [0.0, -0.02, 0.6]], np.float32)
    sta = np.array([[-1, 0, -1], [1, 0, -1], [0, 0, 1.5]], np.float32)
    x = np.concatenate([mov, sta])
    tris = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    group = np.array([0, 1], np.int32)
    im = np.array([1, 1, 1, 0, 0, 0], np.float32)
    depth0 = -float(x[:3, 1].min())
    pairs = oracle.dcd_pairs(tris, x, group=group, allow_self=True)
    assert len(pairs) == 1
    xp = x.copy()
    for _ in range(8):
        c, n = oracle.dcd_narrow(pairs, tris, x)
        x = oracle.dcd_respond_apply(c, n, x, xp, im, 1.01, 0.5, 0.4)
    assert -float(x[:3, 1].min()) < 0.5 * depth0


def test_narrow_symmetric_swap_and_disjoint():
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(200):
        x = rng.normal(scale=0.3, size=(6, 3)).astype(np.float32)
        tris = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
        c, _ = oracle.dcd_narrow(np.array([[0, 1]]), tris, x)
        d, _ = oracle.dcd_narrow(np.array([[1, 0]]), tris, x)
        assert c[0].valid == d[0].valid
        if c[0].valid:
            hits += 1
            assert c[0].C1 == d[0].C2 and c[0].C2 == d[0].C1
    assert hits > 5
    # far apart -> no contact
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
    c, _ = oracle.dcd_narrow(np.array([[0, 1]]), np.array([[0, 1, 2], [3, 4, 5]], np.int32), x)
    assert c[0].valid == 0


def test_pairs_sorted_and_match_sphere_predicate():
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, size=(150, 3)).astype(np.float32)
    tris = rng.integers(0, 150, size=(120, 3)).astype(np.int32)
    p = oracle.dcd_pairs(tris, x, allow_self=True, thickness=0.01)
    assert np.all(np.diff(p[:, 0] * 1000 + p[:, 1]) > 0)
    c, r, _ = oracle.dcd_tri_data(tris, x)
    ref = set()
    for i in range(len(tris)):
        for j in range(i + 1, len(tris)):
            if set(tris[i]) & set(tris[j]):
                continue
            if np.linalg.norm(c[i].astype(np.float64) - c[j]) <= r[i] + r[j] + 0.01 - 1e-6:
                ref.add((i, j))
    got = set(map(tuple, p))
    assert ref <= got and len(got - ref) <= 2   # only float-boundary ties may differ
