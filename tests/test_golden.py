"""Committed fixtures (tests/golden, written by make_golden.py from the
one-thread oracle): the generators still produce the same meshes, and the
oracle still reproduces every fixture bit for bit."""
import json
import sys
from pathlib import Path

import numpy as np

import oracle
from rlmusclev2_amd import sims

GOLD = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLD))
import make_golden as mg  # noqa: E402


def test_cantilever_mesh_unchanged():
    _, w = mg.cantilever_case()
    f = np.load(GOLD / "cantilever_10f.npz")
    assert mg.mesh_digest(w) == str(f["digest"])


def test_oracle_reproduces_cantilever_fixture():
    cfg, w = mg.cantilever_case()
    f = np.load(GOLD / "cantilever_10f.npz")
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=lambda sc: oracle.OracleSim(sc, nthreads=1))
    for fr in range(1, 11):
        ref.step_frame(fr)
    np.testing.assert_array_equal(ref.positions_np(), f["positions"])
    res = ref.residuals()
    np.testing.assert_array_equal(np.array([res[k] for k in f["residual_types"]]), f["residuals"])
    assert ref.count_inverted() == int(f["inverted"])


def test_oracle_reproduces_dcd_pairs_fixture():
    f = np.load(GOLD / "dcd_pairs_small.npz")
    tris, x = mg.dcd_case()
    np.testing.assert_array_equal(tris, f["tris"])
    np.testing.assert_array_equal(x, f["x"])
    pairs = oracle.dcd_pairs(f["tris"], f["x"], allow_self=True, thickness=float(f["thickness"]))
    np.testing.assert_array_equal(np.asarray(pairs, np.int32), f["pairs"])
