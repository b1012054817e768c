"""Host logic of the bone-mesh collider (CPU): the config and class wiring that
the GPU path shares, run on the oracle engine."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import collider, sims, workloads

CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


def _small():
    cfg = json.loads((CFG / "humanv2_mesh_collider.json").read_text())
    cfg["demo"]["collision"]["mesh_res"] = [24, 4, 16]
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=3000, fascia_tets=800, sdf_res=16)
    return cfg, w


def test_capsule_bone_meshes_are_closed_and_move_rigidly():
    cfg, w = _small()
    v, t, toff, vb = workloads.bone_meshes(w, n_around=24, n_cap=4, n_len=16)
    assert toff[0] == 0 and toff[-1] == len(t) and len(toff) == len(w.bones) + 1
    for b in range(len(w.bones)):
        tb = t[toff[b]:toff[b + 1]]
        assert (vb[tb] == b).all()
        a, bb, c = v[tb[:, 0]].astype(np.float64), v[tb[:, 1]].astype(np.float64), v[tb[:, 2]].astype(np.float64)
        assert np.einsum("ij,ij->i", a, np.cross(bb, c)).sum() > 0          # outward: positive volume
    p5 = workloads.bone_mesh_points(w, 5, n_around=24, n_cap=4, n_len=16)
    for b, bn in enumerate(w.bones):
        R, tr = workloads.bone_transform(bn, 5)
        np.testing.assert_allclose(p5[vb == b], v[vb == b] @ R.T + tr, atol=1e-6)


def test_config_collider_mesh_builds_the_scene_and_meter():
    cfg, w = _small()
    sim = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    bm = sim.scene.bone_mesh
    assert sim.scene.sdf is None and bm is not None and bm.resolve
    assert bm.nbones == len(w.bones)
    assert (bm.coll_w[w.is_fascia] == 0).all() and (bm.coll_w[~w.is_fascia] == 1).all()
    sim.step_frame(1)
    d, n, mx = sim.engine.bone_penetration(1e-6)
    assert d.shape == (w.nv,) and n == int((d > 1e-6).sum()) and mx == pytest.approx(float(d.max()))
    bad = dict(cfg, demo={"collision": {"enabled": True, "collider": "bvh"}})
    with pytest.raises(ValueError):
        sims.MultiGridSim(bad, workload=w, engine_cls=oracle.OracleSim)


def test_reference_style_collider_in_collision_solvers():
    """HumanV2.__init__ wiring (humanV2.py:596-622): the collider built after the
    sim, placed in collision_solvers, moved per frame, metered per frame."""
    cfg, w = _small()
    cfg["demo"]["collision"] = {"enabled": False}
    sim = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    src = collider.CapsuleBoneSource(w, n_around=24, n_cap=4, n_len=16)
    col = collider.AnimatedBoneMeshCollider(sim, src, max_dist=0.2, contact_offset=1e-5, max_push=-1.0)
    assert col.n_bone_tris == len(src.tris) // 3
    meter = collider.CollisionMeter(col, None, 1e-6, sim.NV)
    assert not sim.engine.bm_resolve
    sim.collision_solvers = [col]
    assert sim.engine.bm_resolve and sim._py_solvers == []
    col.preload(range(1, 3))
    for f in (1, 2):
        col.set_frame(f)
        sim.step_frame(f)
        meter.measure()
    assert meter.bone_flag.sum() == 0          # the push is the substep's last word
    assert "coll=0(" in meter.summary()
    sim.collision_solvers = []
    assert not sim.engine.bm_resolve
    other = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    with pytest.raises(ValueError):
        other.collision_solvers = [col]
