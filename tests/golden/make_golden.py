"""Writes the golden fixtures of this directory from the CPU oracle (one
thread, so every sum has one order) on seeded synthetic inputs.

    python tests/golden/make_golden.py

* ``cantilever_10f.npz``: the 16x4x4 cantilever (384 tets / 425 verts, seed
  0, ``configs/cantilever.json``) after 10 frames: positions, per-type
  residuals, inverted-tet count, plus a digest of the generated mesh so a
  change of the generator is caught before the numbers are compared.
* ``dcd_pairs_small.npz``: the surface triangles of a small synthetic muscle
  mesh, jittered positions and the exhaustive sorted (i < j) pair list of the
  broadphase emit rules.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from rlmusclev2_amd import sims, workloads  # noqa: E402
from rlmusclev2_amd.collision import surface_tris  # noqa: E402

HERE = Path(__file__).resolve().parent
CFG = ROOT / "rlmusclev2_amd" / "configs"


def mesh_digest(w) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(w.pos0, np.float32).tobytes())
    h.update(np.ascontiguousarray(w.tets, np.int32).tobytes())
    return h.hexdigest()


def cantilever_case():
    cfg = json.loads((CFG / "cantilever.json").read_text())
    w = workloads.cantilever(0)
    return cfg, w


def dcd_case():
    w = workloads.humanv2_scale(4, n_bones=1, per_bone=3, target_tets=3000, fascia_tets=600, sdf_res=16)
    tris = surface_tris(w.tets)
    rng = np.random.default_rng(3)
    x = (w.pos0 + rng.normal(scale=2e-3, size=w.pos0.shape)).astype(np.float32)
    return tris, x


def main():
    cfg, w = cantilever_case()
    ref = sims.MultiGridSim(cfg, workload=w, engine_cls=lambda sc: oracle.OracleSim(sc, nthreads=1))
    for f in range(1, 11):
        ref.step_frame(f)
    res = ref.residuals()
    np.savez_compressed(HERE / "cantilever_10f.npz", digest=np.array(mesh_digest(w)),
                        positions=ref.positions_np().astype(np.float32),
                        residual_types=np.array(sorted(res)), residuals=np.array([res[k] for k in sorted(res)]),
                        inverted=np.array(ref.count_inverted()))
    tris, x = dcd_case()
    pairs = oracle.dcd_pairs(tris, x, allow_self=True, thickness=1e-3)
    np.savez_compressed(HERE / "dcd_pairs_small.npz", tris=tris.astype(np.int32), x=x, thickness=np.float32(1e-3),
                        pairs=np.asarray(pairs, np.int32))
    print("cantilever:", res, "dcd pairs:", len(pairs))


if __name__ == "__main__":
    main()
