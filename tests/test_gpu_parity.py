"""HIP substep vs the CPU oracle on the same seeded inputs (through the C ABI)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import sims, workloads

pytestmark = pytest.mark.gpu
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


def _cfg(name, **over):
    cfg = json.loads((CFG / name).read_text())
    cfg.update(over)
    return cfg


def _bbox_diag(p):
    return float(np.linalg.norm(p.max(0) - p.min(0)))


def _run_pair(cfg, w, frames, patch_verts=256, strategy=0, block_tets=512):
    sim = sims.MultiGridSim(cfg, workload=w, patch_verts=patch_verts, strategy=strategy, block_tets=block_tets)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    for f in range(1, frames + 1):
        print(f"frame {f}", flush=True)
        sim.step_frame(f)
        ref.step_frame(f)
    return sim, ref


def test_cantilever_10_frames_matches_oracle():
    cfg = _cfg("cantilever.json")
    w = workloads.cantilever(0)
    sim, ref = _run_pair(cfg, w, 10)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _bbox_diag(w.pos0)
    err = np.abs(p - q).max()
    print("cantilever max |dx|", err, "tol", tol)
    assert np.all(np.isfinite(p))
    assert err <= tol
    rs, rr = sim.residuals(), ref.residuals()
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])


@pytest.mark.parametrize("strategy,patch_verts", [(0, 64), (0, 256), (1, 256), (2, 256)])
def test_small_humanv2_pins_sdf_matches_oracle(strategy, patch_verts):
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=12000, fascia_tets=3000, sdf_res=48)
    cfg["synthetic"]["sdf_res"] = 48
    sim, ref = _run_pair(cfg, w, 10, patch_verts, strategy)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _bbox_diag(w.pos0)
    err = np.abs(p - q).max()
    print("humanv2-small max |dx|", err, "tol", tol)
    assert err <= tol
    rs, rr = sim.residuals(), ref.residuals()
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])
    assert sim.count_inverted() == ref.count_inverted()


@pytest.mark.parametrize("strategy", [0, 1, 2])
def test_snh_aniso_activation_matches_oracle(strategy):
    """SNH + aniso_arap with an activation ramp (c3 constraint set, small mesh)."""
    cfg = _cfg("humanv2_snh_aniso.json", use_multigrid=False)   # the two-level cycle has its own tests
    w = workloads.humanv2_scale(3, n_bones=1, per_bone=4, target_tets=8000, fascia_tets=1500, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    sim = sims.MultiGridSim(cfg, workload=w, strategy=strategy, block_tets=256)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    for f in range(1, 6):
        a = np.full(w.nt, min(1.0, f / 5.0), np.float32)
        sim.set_fiber_activation(a)
        ref.set_fiber_activation(a)
        sim.step_frame(f)
        ref.step_frame(f)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _bbox_diag(w.pos0)
    err = np.abs(p - q).max()
    print("snh+aniso max |dx|", err, "tol", tol)
    assert err <= tol
    rs, rr = sim.residuals(), ref.residuals()
    assert set(rr) == {"snh", "aniso"}
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])


@pytest.mark.parametrize("strategy", [0, 2])
def test_batch_instances_match_oracle(strategy):
    """uneven instances in one launch, each with its own bone SDF group."""
    cfg = _cfg("batch512.json")
    w = workloads.batch(4, n_inst=6, lo=500, hi=3000, sdf_res=24)
    cfg["synthetic"]["sdf_res"] = 24
    sim = sims.MultiGridSim(cfg, workload=w, patch_verts=128, strategy=strategy, block_tets=512)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    rng = np.random.default_rng(0)
    inst_act = rng.uniform(0, 1, w.info["n_inst"]).astype(np.float32)
    for f in range(1, 4):
        inst_act = np.clip(inst_act + rng.normal(0, 0.1, inst_act.shape), 0, 1).astype(np.float32)
        a = inst_act[w.tet_instance]
        sim.set_fiber_activation(a)
        ref.set_fiber_activation(a)
        sim.step_frame(f)
        ref.step_frame(f)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _bbox_diag(w.pos0[w.vertex_group == 0])
    err = np.abs(p - q).max()
    print("batch max |dx|", err, "tol", tol)
    assert err <= tol


def test_deterministic_and_graph_equals_stream():
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=1, per_bone=4, target_tets=6000, fascia_tets=1000, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    outs = []
    for use_graph, strat in ((True, 0), (True, 0), (False, 0), (True, 1), (True, 1), (True, 2), (True, 2),
                             (False, 2)):
        sim = sims.MultiGridSim(cfg, workload=w, strategy=strat, block_tets=256)
        sim.use_graph = use_graph
        for f in range(1, 4):
            sim.step_frame(f)
        outs.append(sim.positions_np())
    np.testing.assert_array_equal(outs[0], outs[1])   # run-to-run bitwise identical
    np.testing.assert_array_equal(outs[0], outs[2])   # graph replay == stream launches
    np.testing.assert_array_equal(outs[3], outs[4])   # flat strategy: run-to-run identical
    np.testing.assert_allclose(outs[0], outs[3], atol=1e-5)
    np.testing.assert_array_equal(outs[5], outs[6])   # block strategy: run-to-run identical
    np.testing.assert_array_equal(outs[5], outs[7])   # block strategy: graph replay == stream launches
    np.testing.assert_allclose(outs[0], outs[5], atol=1e-5)


@pytest.mark.parametrize("strategy", [1, 2])
def test_phase_api_equals_fused_iteration(strategy):
    """clear / solve_constraints / apply_dP / apply_pintotarget == one fused iteration."""
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=1, per_bone=4, target_tets=6000, fascia_tets=1000, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    a = sims.MultiGridSim(cfg, workload=w)
    b = sims.MultiGridSim(cfg, workload=w, strategy=strategy, block_tets=256)
    for s in (a, b):
        s.load_frame_targets(3)
        s.engine.apply_pintotarget()
        s.engine.integrate()
    a.engine.iterate(1)
    e = b.engine
    e.clear()
    e.solve_constraints()
    e.apply_dP()
    e.apply_pintotarget()
    np.testing.assert_allclose(a.positions_np(), b.positions_np(), atol=1e-7)


def test_sdf_cells_and_push_bit_exact():
    """same positions in -> identical SDF cell indices and pushed positions."""
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=3, target_tets=6000, fascia_tets=1000, sdf_res=40)
    cfg["synthetic"]["sdf_res"] = 40
    sim = sims.MultiGridSim(cfg, workload=w)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    rng = np.random.default_rng(1)
    x = (w.pos0 + rng.normal(scale=5e-3, size=w.pos0.shape)).astype(np.float32)
    for s in (sim, ref):
        s.load_frame_targets(7)
        s.engine.set_state(x, x, np.zeros_like(x))
        s.engine.collide()
    cg, cr = sim.engine.sdf_cells(), ref.engine.sdf_cells()
    assert (cr[:, 0] >= 0).sum() > 100
    np.testing.assert_array_equal(cg, cr)
    np.testing.assert_array_equal(sim.positions_np(), ref.positions_np())


def test_pin_targets_fast_path_and_iteration_timing():
    """per-frame targets go through the on-device scatter after the first full
    upload: pinned vertices land exactly on them; timing leaves the state alone"""
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=6000, fascia_tets=1000, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    sim = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256)
    for f in (1, 5, 9):                 # the first call uploads the table, the rest scatter
        sim.step_frame(f)
        t = workloads.pin_targets(w, f + 1)
        sim.engine.set_pin_targets(t)
        sim.engine.apply_pintotarget()
        p = sim.positions_np()
        np.testing.assert_array_equal(p[w.pin_vids], t)
    x0 = sim.positions_np()
    us = sim.engine.time_iterations(10, 3)
    print("us per iteration", us)
    assert us > 0.0
    np.testing.assert_array_equal(sim.positions_np(), x0)


def test_block_iteration_sequence_equals_single_iterations():
    """xpbd_iterate(n) == n calls of xpbd_iterate(1), bitwise (block strategy,
    odd and even n: the odd count ends with the copy back into X[0])"""
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=6000, fascia_tets=1000, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    a = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256)
    b = sims.MultiGridSim(cfg, workload=w, strategy=2, block_tets=256)
    for s in (a, b):
        s.load_frame_targets(4)
        s.engine.apply_pintotarget()
        s.engine.integrate()
    for n in (5, 4):
        a.engine.iterate(n)
        for _ in range(n):
            b.engine.iterate(1)
        np.testing.assert_array_equal(a.positions_np(), b.positions_np())


@pytest.mark.parametrize("strategy", [1, 2])
@pytest.mark.parametrize("kind", ["inverted", "flat"])
def test_inverted_and_flat_tets_match_oracle(strategy, kind):
    """Tets pushed through themselves to half their height on the other side
    (det F < 0 with a distinct smallest singular value: the SVD fallback and
    the reference's sign convention) or squashed to a tenth of their height
    (det F small and positive), projected once on the GPU and in the oracle
    from the same state.  Not parity cases, in the reference as here: a tet
    flattened exactly (det F = 0 up to rounding: the rotation follows the
    sign of the rounding) or mirrored exactly (F a reflection, all singular
    values 1: every axis is as close, and which one flips is the SVD's
    choice)."""
    cfg = _cfg("cantilever.json")
    w = workloads.cantilever(0)
    sim = sims.MultiGridSim(cfg, workload=w, strategy=strategy, block_tets=128)
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=oracle.OracleSim)
    x = w.pos0.copy()
    t = w.tets
    rng = np.random.default_rng(0)
    pick = rng.choice(len(t), size=12, replace=False)
    for e in pick:
        a, b, c, d = t[e]
        n = np.cross(x[b] - x[a], x[c] - x[a])
        n /= np.linalg.norm(n)
        h = np.dot(x[d] - x[a], n)
        x[d] = x[d] - (1.5 if kind == "inverted" else 0.9) * h * n
    for s in (sim, ref):
        s.engine.set_state(x, x, np.zeros_like(x))
    n_inv = ref.count_inverted()
    print(kind, "inverted before", n_inv, sim.count_inverted())
    assert sim.count_inverted() == n_inv
    assert (n_inv > 0) == (kind == "inverted")
    for s in (sim, ref):
        s.engine.iterate(1)
    p, q = sim.positions_np(), ref.positions_np()
    assert np.all(np.isfinite(p))
    tol = 1e-4 * _bbox_diag(w.pos0)
    err = np.abs(p - q).max()
    print("max |dx| after one iteration", err, "tol", tol, "inverted after", sim.count_inverted(), ref.count_inverted())
    assert err <= tol
    assert sim.count_inverted() == ref.count_inverted()
