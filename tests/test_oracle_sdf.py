"""Bone push-out on trilinear SDF grids (oracle): pinned by closed forms.

Trilinear interpolation reproduces affine fields exactly, so a planar SDF
baked on the grid gives the exact distance and normal; the cell index is
checked against a float64 brute-force floor; the capsule grid is checked
against the analytic capsule distance.  Push rule: solve_mesh_contact_pos_kernel
(src/muscle/collision/sdf/kernel.py:123-170).
"""
import ctypes as C

import numpy as np

import oracle
from rlmusclev2_amd import workloads


def _push(pos, grids, origins, spacing, xf, offset=0.0, max_push=-1.0, coll=None, vgroup=None, goff=None):
    lib = oracle.load()
    pos = np.ascontiguousarray(pos, np.float32).copy()
    n = pos.shape[0]
    coll = np.ones(n, np.float32) if coll is None else np.ascontiguousarray(coll, np.float32)
    cells = np.full((n, 2), -1, np.int32)
    g = np.ascontiguousarray(grids, np.float32)
    o = np.ascontiguousarray(origins, np.float32)
    sp = np.ascontiguousarray(spacing, np.float32)
    x = np.ascontiguousarray(xf, np.float32)
    p = lambda a: None if a is None else C.c_void_p(a.ctypes.data)
    k = lib.orc_sdf_grid_push(n, p(pos), p(coll), g.shape[0], g.shape[1], p(g), p(o), p(sp), p(x),
                              p(vgroup), p(goff), C.c_float(offset), C.c_float(max_push), p(cells))
    return pos, cells, k


def _plane_grid(res, h, origin, n, c):
    t = np.arange(res) * h
    Z, Y, X = np.meshgrid(origin[2] + t, origin[1] + t, origin[0] + t, indexing="ij")
    return (n[0] * X + n[1] * Y + n[2] * Z - c).astype(np.float32)


def test_plane_push_is_exact():
    n = np.array([0.0, 0.0, 1.0])
    res, h = 17, 0.125
    origin = np.array([-1.0, -1.0, -1.0])
    grid = _plane_grid(res, h, origin, n, 0.0)[None]
    xf = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)
    pts = np.array([[0.1, 0.2, -0.3], [0.33, -0.5, 0.25], [0.0, 0.0, -0.01]], np.float32)
    out, cells, k = _push(pts, grid, origin[None], [h], xf, offset=0.05)
    assert k == 2
    np.testing.assert_allclose(out[0], [0.1, 0.2, 0.05], atol=1e-6)
    np.testing.assert_allclose(out[1], pts[1], atol=0)
    np.testing.assert_allclose(out[2], [0.0, 0.0, 0.05], atol=1e-6)
    # cell index = floor((x - origin)/h) in float64 for these well-separated points
    for i in range(3):
        ijk = np.floor((pts[i].astype(np.float64) - origin) / h).astype(int)
        assert cells[i, 1] == (ijk[2] * res + ijk[1]) * res + ijk[0]
        assert cells[i, 0] == 0


def test_rotated_bone_transform():
    """Tilted plane via the bone transform: world = R*local + t."""
    res, h = 21, 0.1
    origin = np.array([-1.0, -1.0, -1.0])
    grid = _plane_grid(res, h, origin, np.array([0.0, 0.0, 1.0]), 0.0)[None]
    ang = 0.4
    R = np.array([[1, 0, 0], [0, np.cos(ang), -np.sin(ang)], [0, np.sin(ang), np.cos(ang)]])
    t = np.array([0.05, -0.02, 0.1])
    xf = np.concatenate([R, t[:, None]], 1).reshape(1, 12)
    nw = R @ [0, 0, 1]
    pts = (np.array([[0.1, 0.2, -0.2]]) @ R.T + t).astype(np.float32)
    out, _, k = _push(pts, grid, origin[None], [h], xf)
    assert k == 1
    phi = float((pts[0] - t) @ nw)
    np.testing.assert_allclose(out[0], pts[0] - phi * nw, atol=2e-6)


def test_capsule_grid_matches_analytic():
    a, b, r = [0.0, 0.0, -0.1], [0.0, 0.0, 0.1], 0.03
    g, o, h = workloads.capsule_sdf_grid(a, b, r, 64, 0.03)
    rng = np.random.default_rng(0)
    pts = rng.uniform(o + 2 * h, o + 61 * h, size=(2000, 3)).astype(np.float32)
    phi = workloads.capsule_phi(pts.astype(np.float64), a, b, r)
    inside = phi < -0.002
    xf = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]], np.float32)
    out, cells, k = _push(pts, g[None], o[None], [h], xf)
    assert k == int((phi < 0).sum()) or abs(k - int((phi < 0).sum())) < 20
    phi_after = workloads.capsule_phi(out[inside].astype(np.float64), a, b, r)
    assert np.all(phi_after > -0.2 * h)            # pushed to the surface within grid error
    assert np.all(cells[:, 0] == 0)


def test_groups_restrict_bones_and_skip_weight_zero():
    res, h = 9, 0.25
    origin = np.array([-1.0, -1.0, -1.0])
    g0 = _plane_grid(res, h, origin, np.array([0.0, 0.0, 1.0]), 0.0)
    g1 = _plane_grid(res, h, origin, np.array([0.0, 0.0, -1.0]), 0.0)
    grids = np.stack([g0, g1])
    xf = np.tile(np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32), (2, 1))
    pts = np.array([[0, 0, -0.1], [0, 0, -0.1], [0, 0, -0.1]], np.float32)
    vgroup = np.array([0, 1, 0], np.int32)
    goff = np.array([0, 1, 2], np.int32)
    coll = np.array([1, 1, 0], np.float32)
    out, cells, k = _push(pts, grids, np.stack([origin, origin]), [h, h], xf, coll=coll, vgroup=vgroup, goff=goff)
    np.testing.assert_allclose(out[0], [0, 0, 0], atol=1e-6)      # bone 0: pushed up to z=0
    np.testing.assert_allclose(out[1], pts[1], atol=0)           # bone 1: phi = +0.1, no contact
    np.testing.assert_allclose(out[2], pts[2], atol=0)           # weight 0: skipped
    assert cells[1, 0] == 1 and cells[2, 0] == -1
