"""The bone-mesh collider's checker and the library's tree, on the CPU.

The oracle's query (brute-force closest point + angle-weighted pseudonormal
sign, restating Warp's mesh_query_point_sign_normal as called by
solve_mesh_contact_pos_kernel, src/muscle/collision/sdf/kernel.py:123-170) is
pinned against geometry with known answers:
  * an L-shaped prism (planar faces, a concave edge): exact distance from its
    rectangles, inside/outside from the L polygon;
  * a thin wedge (a 10 degree convex edge): the sign there needs the edge
    pseudonormal; the same test run with face normals only fails;
  * capsules: distance within the chord error of the tessellation;
  * reversed winding negates the sign; the push lands on contact_offset.
The library's tree (tables, refit, stackless traversal; xpbd_bonemesh.h built
with g++) is then compared with the pinned oracle on the same points.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import workloads

ROOT = Path(__file__).resolve().parents[1]


def prism(poly2d, z0, z1, cap_tris):
    """Closed prism over a CCW polygon (vertex list) with the given cap
    triangulation (CCW index triples); outward winding."""
    n = len(poly2d)
    P = np.asarray(poly2d, np.float64)
    verts = np.concatenate([np.c_[P, np.full(n, z0)], np.c_[P, np.full(n, z1)]])
    tris = []
    for a, b, c in cap_tris:
        tris.append((n + a, n + b, n + c))      # top, +z
        tris.append((a, c, b))                  # bottom, -z
    for i in range(n):
        j = (i + 1) % n
        tris.append((i, j, n + j))
        tris.append((i, n + j, n + i))
    return verts.astype(np.float32), np.asarray(tris, np.int32)


# L polygon (CCW) with (0, 1) on the left side, so that the caps need no T-junction
L_POLY = [(0, 0), (2, 0), (2, 1), (1, 1), (1, 2), (0, 2), (0, 1)]
L_CAP = [(0, 1, 2), (0, 2, 3), (0, 3, 6), (6, 3, 4), (6, 4, 5)]
L_RECTS = [((0, 2), (0, 1)), ((0, 1), (1, 2))]


def l_prism():
    return prism(L_POLY, 0.0, 1.0, L_CAP)


def l_exact(p):
    """signed distance to the L prism from its faces (axis-aligned rectangles)"""
    p = np.asarray(p, np.float64)
    faces = []
    for (x0, x1), (y0, y1) in L_RECTS:           # caps
        for z in (0.0, 1.0):
            faces.append(((x0, x1), (y0, y1), (z, z)))
    walls = [((0, 2), (0, 0)), ((2, 2), (0, 1)), ((1, 2), (1, 1)), ((1, 1), (1, 2)), ((0, 1), (2, 2)),
             ((0, 0), (0, 2))]
    for (x0, x1), (y0, y1) in walls:
        faces.append(((x0, x1), (y0, y1), (0, 1)))
    d = np.full(len(p), np.inf)
    for (x0, x1), (y0, y1), (z0, z1) in faces:
        c = np.stack([np.clip(p[:, 0], x0, x1), np.clip(p[:, 1], y0, y1), np.clip(p[:, 2], z0, z1)], 1)
        d = np.minimum(d, np.linalg.norm(p - c, axis=1))
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    inside = (z > 0) & (z < 1) & (((x > 0) & (x < 2) & (y > 0) & (y < 1)) | ((x > 0) & (x < 1) & (y > 0) & (y < 2)))
    return np.where(inside, -d, d)


def wedge(deg=10.0):
    """triangular prism whose apex edge (at the origin, along z) has the given angle"""
    h = np.tan(np.deg2rad(deg) / 2)
    return prism([(0.0, 0.0), (-1.0, h), (-1.0, -h)], 0.0, 1.0, [(0, 1, 2)]), h


def wedge_inside(p, h):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return (z > 0) & (z < 1) & (x > -1) & (x < 0) & (np.abs(y) < -x * h)


def near_l_features(rng, n):
    """points near the L prism's concave edge, its vertices and edges"""
    feats = np.array([[1, 1, 0.5], [1, 1, 0], [1, 1, 1], [2, 0, 0], [2, 1, 1], [0, 2, 1], [0, 1, 0.5],
                      [1, 2, 0.5], [2, 0.5, 1]], np.float64)
    pick = feats[rng.integers(0, len(feats), n)]
    return (pick + rng.normal(0, 0.05, (n, 3))).astype(np.float32)


def test_l_prism_exact_distance_and_sign():
    v, t = l_prism()
    mq = oracle.MeshQuery(v, t)
    rng = np.random.default_rng(1)
    q = np.concatenate([rng.uniform([-0.5, -0.5, -0.5], [2.5, 2.5, 1.5], (4000, 3)).astype(np.float32),
                        near_l_features(rng, 3000)])
    sd, cp, face, region = mq.query(q, 10.0)
    ref = l_exact(q)
    assert (face >= 0).all()
    np.testing.assert_allclose(np.abs(sd), np.abs(ref), atol=1e-6)
    clear = np.abs(ref) > 1e-5
    assert (np.sign(sd[clear]) == np.sign(ref[clear])).all()
    # every Voronoi region of the triangles was visited
    assert set(np.unique(region)) == set(range(7))


def test_l_prism_max_dist_and_groups():
    v, t = l_prism()
    mq = oracle.MeshQuery(v, t)
    q = np.array([[3.0, 0.5, 0.5], [2.05, 0.5, 0.5]], np.float32)
    sd, _, face, _ = mq.query(q, 0.5)
    assert face[0] == -1 and sd[0] == 0.0            # nothing nearer than max_dist
    assert face[1] >= 0 and abs(sd[1] - 0.05) < 1e-6
    sd2, _, face2, _ = mq.query(q, 1.0 + 1e-3)       # faces at exactly max_dist are not considered
    assert face2[0] >= 0 and abs(sd2[0] - 1.0) < 1e-6


@pytest.mark.parametrize("use_edge_normals", [True, False])
def test_thin_wedge_sign_needs_the_edge_pseudonormal(use_edge_normals):
    (v, t), h = wedge(10.0)
    mq = oracle.MeshQuery(v, t)
    if not use_edge_normals:
        mq.adj = None                                # planted mistake: face normal only on edges
    rng = np.random.default_rng(2)
    # outside, beyond the sharp apex edge, spread over its whole normal cone
    ang = rng.uniform(-np.pi / 2 + 0.1, np.pi / 2 - 0.1, 2000)
    r = rng.uniform(0.01, 0.2, 2000)
    q = np.c_[r * np.cos(ang), r * np.sin(ang), rng.uniform(0.2, 0.8, 2000)].astype(np.float32)
    inside = wedge_inside(q.astype(np.float64), h)
    assert not inside.any()
    sd, _, _, region = mq.query(q, 10.0)
    wrong = int((sd < 0).sum())
    if use_edge_normals:
        assert wrong == 0
        assert (region >= 4).sum() > 500               # most of them meet the apex edge
    else:
        assert wrong > 50                             # the check tells the mistake apart


def test_capsule_distance_within_chord_error():
    n_around, n_cap, n_len = 48, 8, 32
    a, b, r = np.array([0.0, 0, 0]), np.array([0.0, 0, 0.5]), 0.025
    v, t = workloads.capsule_mesh(a, b, r, n_around, n_cap, n_len)
    mq = oracle.MeshQuery(v, t)
    rng = np.random.default_rng(3)
    q = rng.uniform([-0.06, -0.06, -0.06], [0.06, 0.06, 0.56], (8000, 3)).astype(np.float32)
    sd, _, face, _ = mq.query(q, 0.2)
    phi = workloads.capsule_phi(q, a, b, r)
    bound = r * (1 - np.cos(np.pi / n_around) * np.cos(np.pi / (4 * n_cap))) + 1e-6
    assert (face >= 0).all()
    assert np.abs(sd - phi).max() <= bound
    clear = np.abs(phi) > bound
    assert (np.sign(sd[clear]) == np.sign(phi[clear])).all()


def test_reversed_winding_negates_and_adjacency_is_mutual():
    v, t = workloads.capsule_mesh([0, 0, 0], [0.1, 0.2, 0.3], 0.03, 24, 6, 12)
    adj = oracle.mesh_adjacency(t)
    assert (adj >= 0).all()
    for f in range(len(t)):
        for e in range(3):
            g = adj[f, e]
            u, w = t[f, e], t[f, (e + 1) % 3]
            ge = [k for k in range(3) if {t[g, k], t[g, (k + 1) % 3]} == {u, w}]
            assert len(ge) == 1 and adj[g, ge[0]] == f
    rng = np.random.default_rng(4)
    q = rng.uniform(-0.1, 0.45, (3000, 3)).astype(np.float32)
    sd, _, _, _ = oracle.MeshQuery(v, t).query(q, 1.0)
    sd2, _, _, _ = oracle.MeshQuery(v, t[:, [0, 2, 1]]).query(q, 1.0)
    np.testing.assert_allclose(sd2, -sd, atol=1e-9)


def test_push_lands_on_the_offset_surface():
    v, t = workloads.capsule_mesh([0, 0, 0], [0, 0, 0.5], 0.025, 48, 8, 32)
    mq = oracle.MeshQuery(v, t)
    rng = np.random.default_rng(5)
    q = rng.uniform([-0.03, -0.03, -0.02], [0.03, 0.03, 0.52], (3000, 3)).astype(np.float32)
    sd0, _, _, _ = mq.query(q, 0.2)
    pos = q.copy()
    off = 1e-3
    n = mq.push(pos, 0.2, off, -1.0)
    assert n == int((sd0 - off < 0).sum())
    sd1, _, _, _ = mq.query(pos, 0.2)
    moved = sd0 < off
    np.testing.assert_allclose(sd1[moved], off, atol=2e-6)
    assert (pos[~moved] == q[~moved]).all()
    # max_push rejects the deep ones
    pos2 = q.copy()
    mq.push(pos2, 0.2, off, 0.005)
    deep = (off - sd0) > 0.005
    assert (pos2[deep] == q[deep]).all()
    d = mq.depth(q, 0.2)
    np.testing.assert_allclose(d, np.where(sd0 < 0, -sd0, 0.0))


def _tree_checker(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "check_bonemesh_tree"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'rlmusclev2_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "check_bonemesh_tree.cpp"), "-o", str(exe)], check=True)
    return exe


def _run_tree(exe, tmp_path, v, t, toff, q, max_dist, off, max_push):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    nq = len(q)
    with open(inp, "wb") as f:
        np.array([len(v), len(t), len(toff) - 1, nq], np.int32).tofile(f)
        np.ascontiguousarray(v, np.float32).tofile(f)
        np.ascontiguousarray(t, np.int32).tofile(f)
        np.ascontiguousarray(toff, np.int32).tofile(f)
        np.ascontiguousarray(q, np.float32).tofile(f)
        np.array([max_dist, off, max_push], np.float32).tofile(f)
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    _run_tree.closed = r.stdout.split()[-1] == "1"
    raw = np.fromfile(out, np.uint8)
    o = 0

    def take(dt, n):
        nonlocal o
        a = raw[o:o + 4 * n].view(dt)
        o += 4 * n
        return a
    sd = take(np.float32, nq)
    cp = take(np.float32, 3 * nq).reshape(-1, 3)
    face = take(np.int32, nq)
    region = take(np.int32, nq)
    pushed = take(np.float32, 3 * nq).reshape(-1, 3)
    depth = take(np.float32, nq)
    return sd, cp, face, region, pushed, depth


def _scenes():
    rng = np.random.default_rng(6)
    v, t = l_prism()
    yield "L prism", v, t, np.array([0, len(t)]), np.concatenate(
        [rng.uniform([-0.5, -0.5, -0.5], [2.5, 2.5, 1.5], (3000, 3)), near_l_features(rng, 2000)]).astype(np.float32)
    (v, t), _ = wedge(10.0)
    yield "wedge", v, t, np.array([0, len(t)]), rng.uniform([-1.2, -0.3, -0.2], [0.3, 0.3, 1.2],
                                                            (4000, 3)).astype(np.float32)
    # three capsule bones as in the humanV2-scale workload, one sub-mesh each
    bones = [dict(a=np.array([x0, 0, -0.25]), b=np.array([x0, 0, 0.25]), radius=0.025) for x0 in (0.0, 0.9, 1.8)]
    vs, ts, toff, off = [], [], [0], 0
    for bn in bones:
        bv, bt = workloads.capsule_mesh(bn["a"], bn["b"], bn["radius"], 64, 12, 64)
        vs.append(bv)
        ts.append(bt + off)
        toff.append(toff[-1] + len(bt))
        off += len(bv)
    v = np.concatenate(vs).astype(np.float32)
    t = np.concatenate(ts)
    idx = rng.integers(0, len(v), 5000)
    q = (v[idx] + rng.normal(0, 0.01, (5000, 3))).astype(np.float32)
    yield "capsules", v, t, np.asarray(toff), q


@pytest.mark.parametrize("case", ["L prism", "wedge", "capsules"])
def test_library_tree_matches_oracle(tmp_path, case):
    exe = _tree_checker(tmp_path)
    name, v, t, toff, q = next(s for s in _scenes() if s[0] == case)
    max_dist, off, max_push = 0.4, 1e-4, -1.0
    sd, cp, face, region, pushed, depth = _run_tree(exe, tmp_path, v, t, toff, q, max_dist, off, max_push)
    mq = oracle.MeshQuery(v, t, toff)
    osd, ocp, oface, _ = mq.query(q, max_dist)
    assert ((face >= 0) == (oface >= 0)).all()
    hit = oface >= 0
    scale = float(np.abs(v).max())
    np.testing.assert_allclose(np.abs(sd[hit]), np.abs(osd[hit]), atol=2e-6 * scale)
    # the closest point: the oracle's wherever it is unique; at near-ties of two
    # features any point at the minimum distance on the surface will do
    same = np.abs(cp - ocp).max(1) <= 2e-5 * scale
    assert same[hit].mean() > 0.999
    np.testing.assert_allclose(np.linalg.norm(q[hit] - cp[hit], axis=1), np.abs(osd[hit]), atol=2e-6 * scale)
    on, _, _, _ = mq.query(cp[hit], max_dist)
    assert np.abs(on).max() <= 2e-6 * scale
    clear = hit & (np.abs(osd) > 1e-5 * scale)
    assert (np.sign(sd[clear]) == np.sign(osd[clear])).all()
    opos = q.copy()
    mq.push(opos, max_dist, off, max_push)
    # pushed positions: the oracle's, but where the closest point is a near-tie
    # (deep inside a curved bone many surface points are almost equidistant) any
    # push of the right length onto the offset surface
    same = np.abs(pushed - opos).max(1) <= 2e-6 * scale
    assert same.mean() > 0.998
    moved = (np.abs(pushed - q).max(1) > 0) | (np.abs(opos - q).max(1) > 0)
    np.testing.assert_array_equal(moved, (np.abs(opos - q).max(1) > 0))
    np.testing.assert_allclose(np.linalg.norm(pushed - q, axis=1)[moved], np.abs(osd[moved] - off),
                               atol=2e-6 * scale)
    tie = moved & ~same
    if tie.any():
        land, _, _, _ = mq.query(pushed[tie], max_dist)       # outside, at most off from the surface
        assert (land >= -2e-6 * scale).all() and (land <= off + 2e-6 * scale).all()
    np.testing.assert_allclose(depth, mq.depth(q, max_dist), atol=2e-6 * scale)


def test_closed_mesh_detection(tmp_path):
    """the coherent mode's precondition, as the library's table builder sees it"""
    exe = _tree_checker(tmp_path)
    v, t = workloads.capsule_mesh([0, 0, 0], [0, 0, 0.5], 0.025, 24, 4, 12)
    q = np.zeros((1, 3), np.float32)
    cases = {"closed": (t, True), "hole": (t[1:], False),
             "one flipped": (np.r_[t[:1, [0, 2, 1]], t[1:]], False)}
    for name, (tt, want) in cases.items():
        _run_tree(exe, tmp_path, v, tt, np.array([0, len(tt)]), q, 0.2, 0.0, -1.0)
        assert _run_tree.closed == want, name
    (wv, wt), _ = wedge(10.0)
    _run_tree(exe, tmp_path, wv, wt, np.array([0, len(wt)]), q, 0.2, 0.0, -1.0)
    assert _run_tree.closed
