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


def _run_pair(cfg, w, frames, patch_verts=256):
    sim = sims.MultiGridSim(cfg, workload=w, patch_verts=patch_verts)
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


@pytest.mark.parametrize("patch_verts", [64, 256])
def test_small_humanv2_pins_sdf_matches_oracle(patch_verts):
    cfg = _cfg("humanv2_scale.json")
    w = workloads.humanv2_scale(1, n_bones=2, per_bone=4, target_tets=12000, fascia_tets=3000, sdf_res=48)
    cfg["synthetic"]["sdf_res"] = 48
    sim, ref = _run_pair(cfg, w, 10, patch_verts)
    p, q = sim.positions_np(), ref.positions_np()
    tol = 1e-4 * _bbox_diag(w.pos0)
    err = np.abs(p - q).max()
    print("humanv2-small max |dx|", err, "tol", tol)
    assert err <= tol
    rs, rr = sim.residuals(), ref.residuals()
    for k in rr:
        assert abs(rs[k] - rr[k]) <= 0.01 * rr[k] + 1e-12, (k, rs[k], rr[k])
    assert sim.count_inverted() == ref.count_inverted()
