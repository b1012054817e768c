"""The HIP path against the committed fixtures (tests/golden): cantilever
positions after 10 frames within 1e-4 x bbox diagonal and residuals within
1 %, and the DCD pair list bit for bit."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from rlmusclev2_amd import sims, workloads
from rlmusclev2_amd.collision import DcdParams, DcdSpec
from rlmusclev2_amd.engine import XpbdEngine

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"
sys.path.insert(0, str(GOLD))
import make_golden as mg  # noqa: E402


@pytest.mark.parametrize("strategy", [1, 2])
def test_cantilever_matches_fixture(strategy):
    cfg, w = mg.cantilever_case()
    f = np.load(GOLD / "cantilever_10f.npz")
    sim = sims.MultiGridSim(cfg, workload=w, strategy=strategy, block_tets=128)
    for fr in range(1, 11):
        sim.step_frame(fr)
    p, q = sim.positions_np(), f["positions"]
    tol = 1e-4 * float(np.linalg.norm(w.pos0.max(0) - w.pos0.min(0)))
    err = float(np.abs(p - q).max())
    print("max |dx|", err, "tol", tol)
    assert err <= tol
    res = sim.residuals()
    for k, v in zip(f["residual_types"], f["residuals"]):
        assert abs(res[str(k)] - v) <= 0.01 * v + 1e-12, (k, res[str(k)], v)


def test_dcd_pairs_match_fixture():
    f = np.load(GOLD / "dcd_pairs_small.npz")
    cfg = json.loads((CFG / "humanv2_scale.json").read_text())
    cfg["synthetic"]["sdf_res"] = 16
    w = workloads.humanv2_scale(4, n_bones=1, per_bone=3, target_tets=3000, fascia_tets=600, sdf_res=16)
    sc = sims.scene_from_config(cfg, w)
    sc.dcd = DcdSpec(tris=f["tris"], params=DcdParams(thickness=float(f["thickness"])))
    eng = XpbdEngine(sc)
    x = f["x"]
    eng.set_state(x, x, np.zeros_like(x))
    eng.dcd_solve(True)
    np.testing.assert_array_equal(eng.dcd_pairs(), f["pairs"])
