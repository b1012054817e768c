"""ATTACH on the HIP path against the oracle: MultiGridSim's belly attach from
JSON (mg.py:700-734, targets refreshed per frame) and MuscleSim's procedural
attach with non-zero rest lengths (constraint/builder.py:347-398)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import builder, capi, sims, workloads
from rlmusclev2_amd.engine import XpbdEngine

pytestmark = pytest.mark.gpu
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


def _diag(p):
    return float(np.linalg.norm(p.max(0) - p.min(0)))


@pytest.mark.parametrize("strategy", [0, 1, 2])
def test_belly_attach_10_frames_match_oracle(strategy):
    """humanV2.json's belly_attach (k = 1e6, damping 0.1) on a small humanV2
    mesh with tendon-tip pins and bone SDFs"""
    cfg = json.loads((CFG / "humanv2_scale.json").read_text())
    cfg["constraints"].append({"type": "attach", "name": "belly_attach", "enable": True, "stiffness": 1e6,
                               "dampingratio": 0.1})
    cfg["synthetic"]["sdf_res"] = 48
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=12000, fascia_tets=3000, sdf_res=48)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=strategy, block_tets=256)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    assert sim.scene.attach_vids is not None and len(sim.scene.attach_vids) > 1000
    assert sim.engine.stats()["feature_mask"] & 32
    for f in range(1, 11):
        sim.step_frame(f)
        ref.step_frame(f)
        print(f"frame {f} max |dx|", float(np.abs(sim.positions_np() - ref.positions_np()).max()), flush=True)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _diag(w.pos0)
    err = np.abs(p - q).max()
    # the attach pulls the belly along with the bones: it moved away from rest
    moved = float(np.abs(q[sim.scene.attach_vids] - w.pos0[sim.scene.attach_vids]).max())
    print("belly attach max |dx|", err, "tol", tol, "belly moved", moved)
    assert moved > 10 * tol
    assert err <= tol
    rs, rr = sim.residuals(), ref.residuals()
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])


@pytest.mark.parametrize("strategy", [0, 1, 2])
@pytest.mark.parametrize("has_compress", [False, True])
def test_procedural_attach_with_rest_length_matches_oracle(strategy, has_compress):
    """MuscleSim.from_procedural: the cantilever's top face attached to its
    nearest bone point (rest length = that distance, damping 0.05).  From the
    same displaced state (some attached vertices pulled beyond their rest
    length, others pushed inside it, previous positions apart so the damping
    acts), three Jacobi iterations on the GPU and in the oracle."""
    w = workloads.cantilever(3)
    bone = np.array([[0.05, 0.13, 0.05], [0.2, 0.14, 0.05], [0.35, 0.12, 0.05]], np.float32)
    mask = (w.pos0[:, 1] > 0.09).astype(np.float32)
    cc = [{"type": "volume", "stiffness": 1e10, "dampingratio": 0.01},
          {"type": "tetarapsvd", "stiffness": 1e6, "dampingratio": 0.01},
          {"type": "attach", "mask_name": "top", "stiffness": 1e5, "dampingratio": 0.05, "mask_threshold": 0.5}]
    fib = np.tile([1.0, 0.0, 0.0], (w.nt, 1))
    sim = sims.MuscleSim.from_procedural(w.pos0, w.tets, fib, bone, constraint_configs=cc, dts=1 / 24,
                                         num_substeps=5, gravity=-9.8, stopped=w.stopped, engine_cls=oracle.OracleSim,
                                         vertex_attrs={"top": mask})
    sc = sim.scene
    sc.params.has_compress = has_compress
    t = sc.cons
    att = t.type == int(builder.ConstraintType.ATTACH)
    assert att.sum() > 20 and np.all(t.restlength[att] > 0.01)
    t.compressionstiffness[att] = 3e3       # used where an attached vertex is inside its rest length
    eng = XpbdEngine(sc, strategy=strategy, block_tets=128)
    ref = oracle.OracleSim(sc)
    rng = np.random.default_rng(2)
    x = w.pos0.copy()
    src = t.pts[att, 0]
    tgt = t.restvector[att, :3]
    u = (x[src] - tgt) / np.linalg.norm(x[src] - tgt, axis=1, keepdims=True)
    x[src] += u * rng.uniform(-0.4, 0.4, size=(len(src), 1)).astype(np.float32) * t.restlength[att][:, None]
    x += rng.normal(scale=1e-3, size=x.shape).astype(np.float32)
    x[w.stopped != 0] = w.pos0[w.stopped != 0]
    pp = x + rng.normal(scale=2e-3, size=x.shape).astype(np.float32)
    for e in (eng, ref):
        e.set_state(x, pp, np.zeros_like(x))
        e.iterate(3)
    p, q = eng.positions_np(), ref.positions_np()
    tol = 1e-4 * _diag(w.pos0)
    err = np.abs(p - q).max()
    moved = float(np.abs(q[src] - x[src]).max())
    print("procedural attach max |dx|", err, "tol", tol, "attached vertices moved", moved)
    assert moved > 100 * tol
    assert np.all(np.isfinite(p)) and err <= tol


def test_second_attach_on_a_vertex_is_unsupported():
    cfg = json.loads((CFG / "cantilever.json").read_text())
    w = workloads.cantilever(0)
    sc = sims.scene_from_config(cfg, w)
    att = builder.create_attach([5, 5], np.zeros((2, 3)), 1e3, 0.0)
    sc.cons = builder.ConsTable.concat([sc.cons, att]).sorted_by_type().contiguous()
    with pytest.raises(capi.XpbdError) as e:
        XpbdEngine(sc, strategy=2)
    assert e.value.code == capi.XPBD_EUNSUPPORTED
