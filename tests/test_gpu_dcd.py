"""HIP tri-tri DCD vs the oracle: bit-exact sorted pair lists, one DCD pass
(narrow + respond + averaged apply) within float tolerance, and the reference's
penetration-resolution scenario."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from rlmusclev2_amd import sims, workloads
from rlmusclev2_amd.collision import DcdParams, DcdSpec, surface_tris
from rlmusclev2_amd.engine import XpbdEngine

pytestmark = pytest.mark.gpu
CFG = Path(__file__).resolve().parents[1] / "rlmusclev2_amd" / "configs"


def _scene(seed=2, thickness=1e-3, cell=0.0):
    cfg = json.loads((CFG / "humanv2_scale.json").read_text())
    w = workloads.humanv2_scale(seed, n_bones=1, per_bone=6, target_tets=15000, fascia_tets=2000, sdf_res=32)
    cfg["synthetic"]["sdf_res"] = 32
    sc = sims.scene_from_config(cfg, w)
    tris = surface_tris(w.tets)
    sc.dcd = DcdSpec(tris=tris, params=DcdParams(thickness=thickness, cell_size=cell))
    return sc, tris


@pytest.mark.parametrize("cell", [0.0, 0.05])
def test_pairs_bit_exact_vs_bruteforce(cell):
    sc, tris = _scene(cell=cell)
    eng = XpbdEngine(sc)
    rng = np.random.default_rng(0)
    x = sc.pos0 + rng.normal(scale=2e-3, size=sc.pos0.shape).astype(np.float32)
    eng.set_state(x, x, np.zeros_like(x))
    eng.dcd_solve(True)
    got = eng.dcd_pairs()
    ref = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    print("pairs", len(got), len(ref), eng.dcd_stats())
    assert len(ref) > 1000
    np.testing.assert_array_equal(got, ref)


def test_pairs_bit_exact_with_stretched_triangles():
    """a few vertices pulled far out stretch their triangles over several grid
    levels: the pair list still equals brute force and nothing is left out"""
    sc, tris = _scene()
    eng = XpbdEngine(sc)
    rng = np.random.default_rng(1)
    x = sc.pos0 + rng.normal(scale=1e-3, size=sc.pos0.shape).astype(np.float32)
    sv = np.unique(tris)
    pick = rng.choice(sv, size=max(8, sv.size // 100), replace=False)
    x[pick] += rng.normal(scale=0.03, size=(pick.size, 3)).astype(np.float32)
    eng.set_state(x, x, np.zeros_like(x))
    eng.dcd_solve(True)
    got = eng.dcd_pairs()
    ref = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    st = eng.dcd_stats()
    print("pairs", len(got), len(ref), st)
    assert st["left_out_tris"] == 0 and st["truncated_tris"] == 0
    assert len(ref) > 1000
    np.testing.assert_array_equal(got, ref)


def test_dcd_pass_matches_oracle():
    sc, tris = _scene()
    eng = XpbdEngine(sc)
    x = sc.pos0.copy()
    eng.set_state(x, x, np.zeros_like(x))
    eng.dcd_solve(True)
    after = eng.positions_np()
    pairs = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    c, n = oracle.dcd_narrow(pairs, tris, x)
    nvalid = sum(c[i].valid for i in range(n))
    _, _, mass = oracle.precompute_rest(sc.pos0, sc.tets, sc.density)
    im = np.where(mass > 0, 1.0 / np.maximum(mass, 1e-30), 0).astype(np.float32)
    ref = oracle.dcd_respond_apply(c, n, x, x, im, 0.3, 0.5, 0.4)
    st = eng.dcd_stats()
    print("contacts", nvalid, st)
    assert st["n_contacts"] == nvalid
    assert nvalid > 0          # overlapping spindles in this workload
    np.testing.assert_allclose(after, ref, atol=2e-6)


def test_penetration_resolves_on_gpu():
    """movable tilted triangle through a static one (tests/test_udcf_dcd.py:289-332)."""
    from rlmusclev2_amd.builder import ConsTable
    from rlmusclev2_amd.scene import Scene, SimParams
    mov = np.array([[-0.5, -0.02, -0.3], [0.5, 0.04, -0.3], [0.0, -0.02, 0.6]], np.float32)
    sta = np.array([[-1, 0, -1], [1, 0, -1], [0, 0, 1.5]], np.float32)
    x = np.concatenate([mov, sta])
    tris = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    sc = Scene(pos0=x, tets=np.zeros((0, 4), np.int32), density=1.0, cons=ConsTable.empty(),
               params=SimParams(dt=1 / 240), mass=np.array([1, 1, 1, 0, 0, 0], np.float32),
               dcd=DcdSpec(tris=tris, group=np.array([0, 1], np.int32),
                           params=DcdParams(thickness=0.0, multiplier=1.01)))
    eng = XpbdEngine(sc)
    depth0 = -float(x[:3, 1].min())
    for it in range(8):
        eng.dcd_solve(it == 0)
    p = eng.positions_np()
    assert eng.dcd_stats()["n_pairs"] == 1
    assert -float(p[:3, 1].min()) < 0.5 * depth0
    np.testing.assert_array_equal(p[3:], sta)


def _soup(n_tris=2000, box=0.01, edge=0.003, seed=5, max_pairs=1 << 22):
    """n_tris separate small triangles packed in a small box: every search
    finds hundreds of partners, so workgroups overflow their LDS pair stage"""
    from rlmusclev2_amd.builder import ConsTable
    from rlmusclev2_amd.scene import Scene, SimParams
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, box, size=(n_tris, 1, 3))
    x = (c + rng.uniform(-edge / 2, edge / 2, size=(n_tris, 3, 3))).reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * n_tris, dtype=np.int32).reshape(-1, 3)
    sc = Scene(pos0=x, tets=np.zeros((0, 4), np.int32), density=1.0, cons=ConsTable.empty(),
               params=SimParams(dt=1 / 240), mass=np.ones(x.shape[0], np.float32),
               dcd=DcdSpec(tris=tris, params=DcdParams(thickness=1e-3, max_pairs=max_pairs)))
    return sc, tris, x


def test_pairs_bit_exact_when_workgroup_stage_overflows():
    sc, tris, x = _soup()
    eng = XpbdEngine(sc)
    eng.dcd_solve(True)
    got = eng.dcd_pairs()
    ref = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    st = eng.dcd_stats()
    print("pairs", len(got), len(ref), st)
    assert len(ref) > 16 * 2048          # far more than the stages of all workgroups hold on average
    assert st["n_pairs"] == len(ref) and st["truncated_tris"] == 0 and st["left_out_tris"] == 0
    np.testing.assert_array_equal(got, ref)


def test_pairs_beyond_capacity_are_dropped_and_counted():
    cap = 10000
    sc, tris, x = _soup(max_pairs=cap)
    eng = XpbdEngine(sc)
    eng.dcd_solve(True)
    got = eng.dcd_pairs()
    ref = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    st = eng.dcd_stats()
    print("pairs", len(got), len(ref), st)
    assert st["n_pairs"] == len(ref) > cap
    assert st["truncated_tris"] == len(ref) - cap
    assert got.shape == (cap, 2)
    keys = got[:, 0].astype(np.int64) * len(tris) + got[:, 1]
    assert np.all(np.diff(keys) > 0)                       # sorted, no duplicates
    assert np.isin(keys, ref[:, 0].astype(np.int64) * len(tris) + ref[:, 1]).all()
    for _ in range(3):                                     # passes over the kept pairs
        eng.dcd_solve(False)
    assert np.isfinite(eng.positions_np()).all()


def test_crowded_surface_stops_at_the_visit_limit():
    """34,000 triangles packed in 1.5 mm: every search would visit them all
    (more than the 32,768-candidate limit), so the searches stop, the pairs
    beyond max_pairs are dropped, both are counted, and a pass over the kept
    pairs stays finite"""
    cap = 1 << 20
    sc, tris, x = _soup(n_tris=34000, box=0.0015, edge=0.0005, seed=9, max_pairs=cap)
    eng = XpbdEngine(sc)
    eng.dcd_solve(True)
    st = eng.dcd_stats()
    print(st)
    assert st["left_out_tris"] == 0
    assert st["n_pairs"] > cap
    assert st["truncated_tris"] > st["n_pairs"] - cap      # stopped searches on top of the dropped pairs
    eng.dcd_solve(False)
    assert np.isfinite(eng.positions_np()).all()


def test_triangles_with_a_non_finite_vertex_are_left_out():
    """a vertex turned NaN: every triangle on it is left out of the broadphase
    and counted, and the pairs of the others equal brute force over them"""
    sc, tris = _scene()
    eng = XpbdEngine(sc)
    rng = np.random.default_rng(3)
    x = sc.pos0 + rng.normal(scale=2e-3, size=sc.pos0.shape).astype(np.float32)
    bad = int(tris[len(tris) // 2, 0])
    x[bad] = np.nan
    eng.set_state(x, x, np.zeros_like(x))
    eng.dcd_solve(True)
    got = eng.dcd_pairs()
    st = eng.dcd_stats()
    on_bad = np.nonzero((tris == bad).any(axis=1))[0]
    keep = np.setdiff1d(np.arange(len(tris)), on_bad)
    sub = oracle.dcd_pairs(tris[keep], x, allow_self=True, thickness=1e-3)
    ref = np.sort(np.stack([keep[sub[:, 0]], keep[sub[:, 1]]], axis=1), axis=1)
    ref = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    print("left out", st["left_out_tris"], len(on_bad), "pairs", len(got), len(ref))
    assert st["left_out_tris"] == len(on_bad) > 0
    assert not np.isin(got, on_bad).any()
    np.testing.assert_array_equal(got, ref)


def test_frames_with_dcd_equal_the_phase_composition():
    """whole frames (one hipGraph: integrate, iterations, two DCD passes, SDF
    push and velocities fused with the next integrate) equal the same substeps
    composed from the phase calls, bit for bit"""
    sc, tris = _scene()
    a, b = XpbdEngine(sc), XpbdEngine(sc)
    rng = np.random.default_rng(4)
    x = sc.pos0 + rng.normal(scale=1.5e-3, size=sc.pos0.shape).astype(np.float32)
    for e in (a, b):
        e.set_state(x, x, np.zeros_like(x))
    nsub, iters = 4, 5
    for f in range(2):
        print(f"frame {f}", flush=True)
        a.native_frame(nsub, iters)
        for s in range(nsub):
            b.integrate()
            b.iterate(iters)
            for k in range(2):
                b.dcd_solve(s == 0 and k == 0)
            b.collide()
            b.update_velocities()
        # the fused last-pass apply in the post kernel only acts on contacts
        na, nb = a.dcd_stats()["n_contacts"], b.dcd_stats()["n_contacts"]
        print("contacts of the last pass", na, nb, flush=True)
        assert na > 0 and na == nb
    pa, pb = a.positions_np(), b.positions_np()
    assert np.isfinite(pa).all() and a.dcd_stats()["n_pairs"] > 0
    assert float(np.abs(pa - x).max()) > 1e-4
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(a.velocities_np(), b.velocities_np())
