"""The benchmark configurations themselves on the GPU: the humanV2-scale scene
at full size against the oracle, a 16-instance batch drawn from the
512-instance distribution against the oracle, activation by instance equal
to activation by tet, and the warm-start schedule with self-collision DCD
(humanV2-style mesh and the 64-instance uneven batch) against plain + DCD
frames from the same start."""
import json
import os
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import sims, warmstart as ws, workloads

pytestmark = pytest.mark.gpu
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"
THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), len(os.sched_getaffinity(0))))


def _cfg(name, **over):
    cfg = json.loads((CFG / name).read_text())
    cfg.update(over)
    return cfg


def _diag(p):
    return float(np.linalg.norm(p.max(0) - p.min(0)))


def _check(sim, ref, tol_pos, label):
    p, q = sim.positions_np(), ref.positions_np()
    err = float(np.abs(p - q).max())
    rs, rr = sim.residuals(), ref.residuals()
    print(label, "max |dx|", err, "tol", tol_pos, "residuals", rs, rr, flush=True)
    assert np.all(np.isfinite(p)) and err <= tol_pos
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])


def test_humanv2_scale_full_size_matches_oracle():
    """534,882 tets, the bench's own config (ARAP-SVD + volume with damping,
    tendon-tip pins, 4 x 128^3 bone SDFs, 10 x 10), two frames"""
    cfg = _cfg("humanv2_scale.json")
    w = sims.build_workload(cfg)
    assert w.nt > 500_000
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=lambda sc: oracle.OracleSim(sc, nthreads=THREADS))
    for f in (1, 2):
        print(f"frame {f}", flush=True)
        sim.step_frame(f)
        ref.step_frame(f)
    _check(sim, ref, 1e-4 * _diag(w.pos0), "humanv2 full size")
    assert sim.count_inverted() == ref.count_inverted()


def test_sixteen_instances_from_the_batch_distribution_match_oracle():
    """16 instances with tet counts log-uniform in [2k, 40k] (the 512-instance
    configuration's distribution), ARAP-SVD + volume + aniso, an activation
    walk per instance (by instance on the GPU, by tet in the oracle), three
    frames"""
    cfg = _cfg("batch512.json")
    s = cfg["synthetic"]
    w = workloads.batch(7, n_inst=16, lo=s["min_tets"], hi=s["max_tets"], sdf_res=s["sdf_res"])
    sizes = np.bincount(w.tet_instance)
    print("instance tets", sizes.min(), sizes.max(), w.nt, flush=True)
    assert sizes.min() >= 1500 and sizes.max() >= 15000
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=lambda sc: oracle.OracleSim(sc, nthreads=THREADS))
    walk = ws.ActivationWalk(w.tet_instance, np.random.default_rng(3))
    sim.set_activation_groups(w.tet_instance)
    for f in range(1, 4):
        a = walk.next_groups()
        sim.set_group_activation(a)
        ref.set_fiber_activation(a[w.tet_instance])
        sim.step_frame(f)
        ref.step_frame(f)
    p, q = sim.positions_np(), ref.positions_np()
    for i in range(16):   # each instance against 1e-4 x its own bbox diagonal
        sel = w.vertex_group == i
        err, tol = float(np.abs(p[sel] - q[sel]).max()), 1e-4 * _diag(w.pos0[sel])
        print("instance", i, "verts", int(sel.sum()), "max |dx|", err, "tol", tol, flush=True)
        assert err <= tol
    _check(sim, ref, np.inf, "16 instances")


def test_group_activation_equals_per_tet_activation():
    cfg = _cfg("batch512.json")
    w = workloads.batch(4, n_inst=6, lo=500, hi=3000, sdf_res=24)
    cfg["synthetic"]["sdf_res"] = 24
    a = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256)
    b = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256)
    a.set_activation_groups(w.tet_instance)
    rng = np.random.default_rng(0)
    for f in range(1, 4):
        act = rng.uniform(0, 1, 6).astype(np.float32)
        a.set_group_activation(act)
        b.set_fiber_activation(act[w.tet_instance])
        a.step_frame(f)
        b.step_frame(f)
    np.testing.assert_array_equal(a.positions_np(), b.positions_np())


def _warm_vs_plain(cfg, w, frames, tol=0.03, contacts=True):
    """warm frames (reference solve every third substep, K = 4 iterations in
    between, no network) with DCD, each audited against the plain 10-iteration
    frame from the same start: a frame either ends within tol of the plain
    residual for every type or falls back, and a fallback frame equals the
    plain + DCD frame from the same start bit for bit"""
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    plain = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256, block_threads=256)
    assert sim.scene.dcd is not None
    for f in range(1, 3):
        sim.step_frame(f)
    st = ws.WarmStartStepper(sim, k_iters=4, audit_every=1, tol=tol, ref_every=3, use_network=False)
    passed = 0
    for f in range(3, 3 + frames):
        plain.engine.set_state(*sim.engine.get_state())
        info = st.step_frame(f)
        plain.step_frame(f)
        d = sim.engine.dcd_stats()
        print(f, "fallback", info["fallback"], info["residual"], info["plain"], "dcd", d, flush=True)
        assert d["n_pairs"] > 0 and (d["n_contacts"] > 0 or not contacts)
        assert set(info["plain"]) == set(plain.residuals())
        if info["fallback"]:
            np.testing.assert_array_equal(sim.positions_np(), plain.positions_np())
        else:
            passed += 1
            for k, v in info["plain"].items():
                assert info["residual"][k] <= v * (1 + tol)
    print("frames within tolerance", passed, "of", frames)
    assert 2 * passed >= frames
    assert np.all(np.isfinite(sim.positions_np()))


def test_warm_schedule_with_dcd_reaches_plain_residual_humanv2_mesh():
    cfg = _cfg("humanv2_c5.json")
    cfg["synthetic"]["sdf_res"] = 32
    w = workloads.humanv2_scale(2, n_bones=1, per_bone=6, target_tets=15000, fascia_tets=2000, sdf_res=32)
    _warm_vs_plain(cfg, w, 6)


def test_warm_schedule_with_dcd_reaches_plain_residual_64_instances():
    """the c5 64-instance uneven batch itself"""
    cfg = _cfg("batch64_c5.json")
    w = sims.build_workload(cfg)
    assert w.info["n_inst"] == 64
    _warm_vs_plain(cfg, w, 6, contacts=False)   # separate instances: pairs, contacts only if one folds
