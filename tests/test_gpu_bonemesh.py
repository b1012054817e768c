"""Bone-mesh collider on the GPU (BVH refit + stackless closest-point query)
against the oracle's brute force (orc_mesh_query), through the C ABI.

Closest points are compared where they are unique; at near-ties (a vertex deep
inside a curved bone sees many almost equidistant surface points) any push of
the right length onto the offset surface passes.
"""
import dataclasses
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import collider, sims, workloads

pytestmark = pytest.mark.gpu
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


def _cfg(**coll):
    cfg = json.loads((CFG / "humanv2_mesh_collider.json").read_text())
    cfg["demo"]["collision"].update(coll)
    return cfg


def _small(seed=1):
    return workloads.humanv2_scale(seed, n_bones=2, per_bone=4, target_tets=12000, fascia_tets=3000, sdf_res=16)


def _perturbed(w, rng, scale=0.01):
    """rest positions pulled towards the bone axes, so that many vertices are inside"""
    x = w.pos0.astype(np.float64).copy()
    for b, bn in enumerate(w.bones):
        sel = w.vertex_group == b
        a, c = np.asarray(bn["a"]), np.asarray(bn["b"])
        u = (c - a) / np.linalg.norm(c - a)
        r = x[sel] - a
        radial = r - np.outer(r @ u, u)
        x[sel] -= 0.5 * radial * rng.uniform(0.0, 1.0, (sel.sum(), 1))
    return (x + rng.normal(0, scale * 0.1, x.shape)).astype(np.float32)


def _compare_push(p_gpu, p_orc, q, osd, off, scale, mq, max_dist):
    same = np.abs(p_gpu - p_orc).max(1) <= 2e-6 * scale
    assert same.mean() > 0.998, same.mean()
    moved_o = np.abs(p_orc - q).max(1) > 0
    moved_g = np.abs(p_gpu - q).max(1) > 0
    np.testing.assert_array_equal(moved_g, moved_o)
    np.testing.assert_allclose(np.linalg.norm(p_gpu - q, axis=1)[moved_o], np.abs(osd[moved_o] - off),
                               atol=2e-6 * scale)
    tie = moved_o & ~same
    if tie.any():
        land, _, _, _ = mq.query(p_gpu[tie], max_dist)
        assert (land >= -2e-6 * scale).all() and (land <= off + 2e-6 * scale).all()


@pytest.mark.parametrize("frame", [1, 7])
def test_push_and_meter_match_oracle(frame):
    cfg = _cfg()
    w = _small()
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    bm = sim.scene.bone_mesh
    pts = workloads.bone_mesh_points(w, frame)
    sim.engine.set_bone_mesh_points(pts)
    rng = np.random.default_rng(frame)
    q = _perturbed(w, rng)
    sim.engine.set_state(pos=q)
    mq = oracle.MeshQuery(pts, bm.tris, bm.tri_off)
    print(f"bone mesh {len(bm.tris)} triangles, {w.nv} vertices, frame {frame}", flush=True)
    # the meter first (it does not move anything)
    d, n, mx = sim.engine.bone_penetration(1e-6)
    od = mq.depth(q, bm.max_dist, bm.meter_w)
    scale = float(np.abs(q).max())
    np.testing.assert_allclose(d, od, atol=2e-6 * scale)
    assert abs(n - int((od > 1e-6).sum())) <= int(0.001 * w.nv) + 1
    assert mx == pytest.approx(float(od.max()), abs=2e-6 * scale)
    assert int((od > 1e-3).sum()) > 0.03 * w.nv    # the case has real penetration
    sim.engine.bone_mesh_push()
    p = sim.positions_np()
    osd, _, _, _ = mq.query(q, bm.max_dist)
    po = q.copy()
    mq.push(po, bm.max_dist, bm.contact_offset, bm.max_push, bm.coll_w)
    _compare_push(p, po, q, osd, bm.contact_offset, scale, mq, bm.max_dist)
    assert (p[w.is_fascia] == q[w.is_fascia]).all()          # fascia is never pushed


def test_max_push_matches_oracle():
    """max_push rejects deep contacts (the sign-flip guard, kernel.py:142-154)"""
    cfg = _cfg(max_push=0.004, contact_offset=1e-4)
    w = _small(4)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    bm = sim.scene.bone_mesh
    rng = np.random.default_rng(3)
    q = _perturbed(w, rng)
    sim.engine.set_state(pos=q)
    sim.engine.bone_mesh_push()
    p = sim.positions_np()
    mq = oracle.MeshQuery(bm.verts, bm.tris, bm.tri_off)
    osd, _, _, _ = mq.query(q, bm.max_dist)
    po = q.copy()
    mq.push(po, bm.max_dist, bm.contact_offset, bm.max_push, bm.coll_w)
    deep = (bm.contact_offset - osd) > bm.max_push
    assert int(deep.sum()) > 0.01 * w.nv              # rejected deep contacts exist
    assert (p[deep] == q[deep]).all()
    _compare_push(p, po, q, osd, bm.contact_offset, float(np.abs(q).max()), mq, bm.max_dist)


def test_vertex_groups_select_the_sub_meshes():
    """each vertex tests only the bones of its group: with the groups swapped the
    vertices pulled into their own bone are outside the one they test"""
    cfg = _cfg()
    w = _small(5)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    bm = sim.scene.bone_mesh
    q = _perturbed(w, np.random.default_rng(5))
    sim.engine.set_state(pos=q)
    mq = oracle.MeshQuery(bm.verts, bm.tris, bm.tri_off)
    goff = np.arange(len(w.bones) + 1, dtype=np.int32)
    for swap in (False, True):
        vg = w.vertex_group.astype(np.int32)
        if swap:
            vg = (vg + 1) % len(w.bones)
        spec = dataclasses.replace(bm, vertex_group=vg, group_bone_off=goff, max_dist=2.0)
        sim.engine.set_bone_mesh(spec)
        d, n, _ = sim.engine.bone_penetration(1e-6)
        od = mq.depth(q, 2.0, bm.meter_w, vg, goff)
        np.testing.assert_allclose(d, od, atol=2e-6 * float(np.abs(q).max()))
        print("swap", swap, "inside", n, flush=True)
        assert (n == 0) if swap else (n > 0.03 * w.nv)


def test_frames_with_mesh_collider_match_oracle():
    cfg = _cfg(mesh_res=[64, 12, 48])
    w = _small(2)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    for f in range(1, 9):
        print(f"frame {f}", flush=True)
        sim.step_frame(f)
        ref.step_frame(f)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * float(np.linalg.norm(w.pos0.max(0) - w.pos0.min(0)))
    err = float(np.abs(p - q).max())
    print("mesh collider frames max |dx|", err, "tol", tol)
    assert err <= tol
    d, n, _ = sim.engine.bone_penetration(1e-6)
    dr, nr, _ = ref.engine.bone_penetration(1e-6)
    assert n == 0 and nr == 0                    # the push is the substep's last word


def test_graph_equals_stream_and_reference_style_collider():
    cfg = _cfg(mesh_res=[48, 8, 32])
    w = _small(3)
    a = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    b = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    b.use_graph = False
    # the reference's wiring: no collider in the config, a collider object in collision_solvers
    c_cfg = _cfg(mesh_res=[48, 8, 32])
    c_cfg["demo"]["collision"] = {"enabled": False}
    c = sims.MultiGridSim(c_cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    src = collider.CapsuleBoneSource(w, n_around=48, n_cap=8, n_len=32)
    col = collider.AnimatedBoneMeshCollider(c, src, max_dist=0.2, contact_offset=1e-5, max_push=-1.0)
    c.collision_solvers = [col]
    meter = collider.CollisionMeter(col, None, 1e-6, c.NV)
    for f in range(1, 5):
        print(f"frame {f}", flush=True)
        a.step_frame(f)
        b.step_frame(f)
        col.set_frame(f)
        c.step_frame(f)
        meter.measure()
        assert meter.bone_flag.sum() == 0
    pa, pb, pc = a.positions_np(), b.positions_np(), c.positions_np()
    assert np.array_equal(pa, pb)
    assert np.array_equal(pa, pc)


def test_full_size_push_sample_and_meter():
    """humanV2-scale mesh (534,882 tets) against the skeleton mesh (97,536
    triangles): one frame, then no metered vertex is inside; a pushed perturbed
    state agrees with the oracle on a sample of 3000 vertices."""
    cfg = _cfg()
    w = sims.build_workload(cfg)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    bm = sim.scene.bone_mesh
    print(f"full size: {w.nt} tets, {len(bm.tris)} bone triangles", flush=True)
    sim.step_frame(1)
    d, n, mx = sim.engine.bone_penetration(1e-6)
    assert n == 0, (n, mx)
    rng = np.random.default_rng(11)
    q = _perturbed(w, rng)
    sim.engine.set_state(pos=q)
    pts = workloads.bone_mesh_points(w, 1)
    mq = oracle.MeshQuery(pts, bm.tris, bm.tri_off)
    d, n, mx = sim.engine.bone_penetration(1e-6)
    idx = rng.choice(w.nv, 3000, replace=False)
    od = mq.depth(q[idx], bm.max_dist, bm.meter_w[idx], nthreads=16)
    scale = float(np.abs(q).max())
    np.testing.assert_allclose(d[idx], od, atol=2e-6 * scale)
    sim.engine.bone_mesh_push()
    p = sim.positions_np()
    osd, _, _, _ = mq.query(q[idx], bm.max_dist, nthreads=16)
    po = q[idx].copy()
    mq.push(po, bm.max_dist, bm.contact_offset, bm.max_push, bm.coll_w[idx], nthreads=16)
    _compare_push(p[idx], po, q[idx], osd, bm.contact_offset, scale, mq, bm.max_dist)
    d2, n2, _ = sim.engine.bone_penetration(1e-6)
    assert n2 == 0


def test_coherent_queries_equal_querying_every_vertex():
    """the certificate skips only queries whose outcome it proves: frames, the meter
    and a state set from outside (vertices teleported into the bones) give the
    same bits with and without it"""
    w = _small(6)
    runs = {}
    for coh in (0, 2):
        cfg = _cfg(coherent=coh, mesh_res=[64, 12, 48])
        sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
        meters = []
        for f in range(1, 9):
            print(f"coherent {coh} frame {f}", flush=True)
            if f == 5:
                sim.engine.set_state(pos=_perturbed(w, np.random.default_rng(9)))
            sim.step_frame(f)
            meters.append(sim.engine.bone_penetration(1e-6))
        runs[coh] = (sim.positions_np(), meters)
    p0, m0 = runs[0]
    p2, m2 = runs[2]
    assert np.array_equal(p0, p2)
    for (d0, n0, x0), (d2, n2, x2) in zip(m0, m2):
        assert np.array_equal(d0, d2) and n0 == n2 and x0 == x2


def test_coherent_mode_needs_a_closed_mesh():
    from rlmusclev2_amd.capi import XpbdError
    cfg = _cfg(mesh_res=[24, 4, 16])
    w = _small(7)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    bm = sim.scene.bone_mesh
    open_mesh = dataclasses.replace(bm, tris=bm.tris[1:], tri_off=np.r_[0, bm.tri_off[1:] - 1].astype(np.int32),
                                    coherent=1)
    with pytest.raises(XpbdError, match="closed"):
        sim.engine.set_bone_mesh(open_mesh)
    sim.engine.set_bone_mesh(dataclasses.replace(open_mesh, coherent=2))      # automatic: off on an open mesh
    q = _perturbed(w, np.random.default_rng(7))
    sim.engine.set_state(pos=q)
    d, n, _ = sim.engine.bone_penetration(1e-6)
    mq = oracle.MeshQuery(bm.verts, open_mesh.tris, open_mesh.tri_off)
    np.testing.assert_allclose(d, mq.depth(q, bm.max_dist, bm.meter_w), atol=2e-6 * float(np.abs(q).max()))
