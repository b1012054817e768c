"""The reference's asset-driven inputs on the host: a humanV2-style mesh written as
restpos.geo + cons_<frame>.geo + bone_<frame>.geo files (the reference's assets are
git-LFS pointers, so the files are synthetic), read back by sims.build_workload;
the pin region rules of muscleToBonePin, the belly-attach set, the per-frame goals
of both target modes against closed forms, the host closest-point bind against the
oracle's brute force, and whole frames of the asset-driven sim on the oracle."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import assets, geo, sims, workloads

CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


@pytest.fixture(scope="module")
def small():
    return workloads.humanv2_scale(1, n_bones=2, per_bone=3, target_tets=5000, fascia_tets=800, sdf_res=32)


def _write_assets(root: Path, w, frames=(1, 2, 3), subset=None):
    """restpos.geo (P, tets, materialW, tendonmask, pinmask), cons_<f>.geo with pts and
    target_pos (the synthetic bone-follow goals moved by a frame-dependent offset), and
    bone_<f>.geo (the capsule skeleton) for the frames"""
    (root / "cons").mkdir(parents=True, exist_ok=True)
    (root / "bone").mkdir(exist_ok=True)
    nv = w.nv
    pinmask = np.zeros(nv, np.float32)
    pinmask[w.pin_vids] = 1.0
    geo.write_geo(root / "restpos.geo", w.pos0, w.tets,
                  point_attrs={"materialW": w.v_fiber_dir, "tendonmask": w.v_tendonmask, "pinmask": pinmask})
    pts = np.arange(nv, dtype=np.int32) if subset is None else np.asarray(subset, np.int32)
    rng = np.random.default_rng(4)
    bias = rng.uniform(-1e-3, 1e-3, (len(pts), 3)).astype(np.float32)     # target_pos != rest
    for f in frames:
        tgt = (workloads.follow_targets(w, f, pts) + bias).astype(np.float32)
        geo.write_geo(root / "cons" / f"cons_{f}.geo", tgt, point_attrs={"pts": pts, "target_pos": tgt})
    src_res = dict(n_around=24, n_cap=4, n_len=24)
    verts, tris, _, _ = workloads.bone_meshes(w, **src_res)
    for f in frames:
        geo.write_geo(root / "bone" / f"bone_{f}.geo", workloads.bone_mesh_points(w, f, **src_res), tris[:, [0, 2, 1]])
    return pts, bias


def _cfg(**over):
    cfg = json.loads((CFG / "humanv2_scale.json").read_text())
    del cfg["synthetic"]
    cfg.update(dict(rest_path="restpos.geo", cons_dir="cons", target_mode="points", start_frame=1))
    cfg["demo"]["collision"]["enabled"] = False
    cfg.update(over)
    return cfg


def test_rest_and_cons_round_trip(tmp_path, small):
    pts, bias = _write_assets(tmp_path, small)
    w = sims.build_workload(_cfg(), base=tmp_path)
    np.testing.assert_array_equal(w.pos0, small.pos0.astype(np.float32))
    np.testing.assert_array_equal(w.tets, small.tets.astype(np.int32))
    np.testing.assert_array_equal(w.v_fiber_dir, small.v_fiber_dir.astype(np.float32))
    np.testing.assert_array_equal(w.pin_vids, pts)
    t = w.info["assets"]["targets"]
    # rest_offset = rest[pts] - target_pos(start) and goal(f) = target_pos(f) + rest_offset
    tp1 = geo.GeoFile(str(tmp_path / "cons" / "cons_1.geo")).target_pos
    np.testing.assert_array_equal(t.rest_offset, (w.pos0[pts] - tp1).astype(np.float32))
    np.testing.assert_array_equal(t.goals(1), (tp1 + t.rest_offset).astype(np.float32))
    tp3 = geo.GeoFile(str(tmp_path / "cons" / "cons_3.geo")).target_pos
    np.testing.assert_array_equal(t.goals(3), (tp3 + t.rest_offset).astype(np.float32))
    # the goal at the start frame is the rest position (to FP32 rounding)
    np.testing.assert_allclose(t.goals(1), w.pos0[pts], atol=1e-6)


def test_pin_regions_and_belly_attach(tmp_path, small):
    pts, _ = _write_assets(tmp_path, small)
    w = sims.build_workload(_cfg(), base=tmp_path)
    tm, pm = w.info["assets"]["point_attrs"]["tendonmask"], w.info["assets"]["point_attrs"]["pinmask"]
    a = assets.pin_active({"region": "tendonmask", "tendon_pin_threshold": 0.5}, w)
    np.testing.assert_array_equal(a, (tm[pts] > 0.5).astype(np.int32))
    np.testing.assert_array_equal(assets.pin_active({"region": "pinmask"}, w), (pm[pts] > 0.5).astype(np.int32))
    assert assets.pin_active({"region": "full"}, w).all()
    assert not assets.pin_active(None, w).any()
    with pytest.raises(ValueError):
        assets.pin_active({"region": "tendonmask"}, w)           # no threshold
    with pytest.raises(ValueError):
        assets.pin_active({"region": "tips"}, w)
    # the scene: hard pins by pinmask, the other constrained vertices get the belly attach
    cfg = _cfg()
    cfg["constraints"] = [c for c in cfg["constraints"] if c["type"] != "pin"] + [
        {"type": "pin", "name": "muscleToBonePin", "enable": True, "region": "pinmask"},
        {"type": "attach", "name": "belly_attach", "enable": True, "stiffness": 1e6, "dampingratio": 0.1}]
    sc = sims.scene_from_config(cfg, w)
    np.testing.assert_array_equal(sc.pin_vids, pts)
    np.testing.assert_array_equal(sc.pin_active, (pm[pts] > 0.5).astype(np.int32))
    np.testing.assert_array_equal(sc.attach_vids, pts[pm[pts] <= 0.5])


def test_pinmask_absent_means_no_hard_pins(tmp_path, small):
    _write_assets(tmp_path, small)
    geo.write_geo(tmp_path / "restpos.geo", small.pos0, small.tets,
                  point_attrs={"materialW": small.v_fiber_dir, "tendonmask": small.v_tendonmask})
    w = sims.build_workload(_cfg(), base=tmp_path)
    assert not assets.pin_active({"region": "pinmask"}, w).any()
    with pytest.raises(NotImplementedError):
        sims.build_workload(_cfg(rest_path="muscle.usd"), base=tmp_path)
    with pytest.raises(NotImplementedError):
        sims.build_workload(_cfg(target_mode="lbs"), base=tmp_path)


def test_mesh_bind_matches_oracle_brute_force(small):
    verts, tris, _, _ = workloads.bone_meshes(small, n_around=24, n_cap=4, n_len=24)
    rng = np.random.default_rng(0)
    q = small.pos0[rng.choice(small.nv, 400, replace=False)] + rng.normal(0, 1e-3, (400, 3))
    face, w, gap = assets.mesh_bind(verts, tris, q, 0.05)
    rf, rw, rg = oracle.fascia_bind(verts, tris, q, 0.05)
    np.testing.assert_array_equal(face >= 0, rf >= 0)
    ok = face >= 0
    assert ok.sum() > 100
    cp = np.einsum("nk,nkd->nd", w[ok].astype(np.float64), verts[tris[face[ok]]].astype(np.float64))
    rcp = np.einsum("nk,nkd->nd", rw[ok], verts[tris[rf[ok]]].astype(np.float64))
    # the same closest point (faces may differ only at ties of equal distance)
    np.testing.assert_allclose(cp, rcp, atol=2e-6)
    np.testing.assert_allclose(np.abs(gap[ok]), np.abs(rg[ok]), atol=2e-6)


def test_bone_surface_targets_follow_rigid_motion(tmp_path, small):
    """target_mode "mesh": a rigid motion of the whole skeleton moves every bound goal
    rigidly, and at the start frame the goals are the rest positions"""
    _write_assets(tmp_path, small, frames=(1,))
    b1 = geo.GeoFile(str(tmp_path / "bone" / "bone_1.geo"))
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]], np.float32)
    tr = np.array([0.01, -0.02, 0.03], np.float32)
    moved = (b1.positions @ R.T + tr).astype(np.float32)
    geo.write_geo(tmp_path / "bone" / "bone_2.geo", moved, b1.tet_indices)
    w = sims.build_workload(_cfg(target_mode="mesh", bone_dir="bone", slide_range=0.05), base=tmp_path)
    t = w.info["assets"]["targets"]
    assert t.ok.mean() > 0.9
    g1, g2 = t.goals(1), t.goals(2)
    ok = t.ok
    np.testing.assert_allclose(g1[ok], w.pos0[ok], atol=2e-6)
    np.testing.assert_allclose(g2[ok], (w.pos0[ok] @ R.T + tr), atol=1e-5)
    with pytest.raises(NotImplementedError):
        sims.build_workload(_cfg(target_mode="mesh", bone_dir="bone", use_sliding=True), base=tmp_path)


def test_asset_frames_on_the_oracle(tmp_path, small):
    """two frames of the asset-driven sim (pinmask hard pins, belly attach, ARAP +
    volume, the .geo skeleton in the exact-mesh collider as a penetration meter) on
    the oracle: hard-pinned vertices sit on their goals, the belly moved toward its
    goals"""
    pts, _ = _write_assets(tmp_path, small)
    cfg = _cfg(bone_dir="bone")
    cfg["constraints"] = [c for c in cfg["constraints"] if c["type"] != "pin"] + [
        {"type": "pin", "name": "muscleToBonePin", "enable": True, "region": "pinmask"},
        {"type": "attach", "name": "belly_attach", "enable": True, "stiffness": 1e6, "dampingratio": 0.1}]
    cfg["demo"]["collision"].update(enabled=False, collider="mesh")
    w = sims.build_workload(cfg, base=tmp_path)
    sim = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    assert sim.scene.bone_mesh is not None and not sim.scene.bone_mesh.resolve
    for f in (2, 3):
        sim.step_frame(f)
    p = sim.positions_np()
    goals = w.info["assets"]["targets"].goals(3)
    act = sim.scene.pin_active != 0
    assert np.isfinite(p).all()
    np.testing.assert_array_equal(p[pts[act]], goals[act])
    d0 = np.abs(w.pos0[pts[~act]] - goals[~act]).max()
    d1 = np.abs(p[pts[~act]] - goals[~act]).max()
    print("belly distance to goals", d0, "->", d1)
    assert d1 < d0
