"""Launch-shape autotuner for the fused Jacobi projection.

Candidates: the "patch" strategy (LDS-staged vertex table, one launch per
iteration) at 64/128/256 owned vertices per workgroup, which trades workgroup
count and occupancy against tet replication across patch borders, and the
"flat" strategy (one thread per tet, then one per vertex; two launches per
iteration, no replication).  Each candidate layout is built for the actual
mesh, a few iterations are timed with HIP events around the projection
launches, and the fastest is kept.  Results are cached per mesh signature in
``rlmusclev2_amd/_tune/cache.json``.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

CACHE = Path(__file__).resolve().parent / "_tune" / "cache.json"
CANDIDATES = ((1, 256), (0, 64), (0, 128), (0, 256))   # (strategy, patch_verts)


def signature(cfg: dict, w) -> str:
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(w.tets).tobytes()[:1 << 20])
    h.update(json.dumps([cfg.get("constraints"), w.nt, w.nv], sort_keys=True, default=str).encode())
    return h.hexdigest()[:16]


def _load():
    try:
        return json.loads(CACHE.read_text())
    except Exception:
        return {}


def time_candidates(cfg: dict, w, candidates=CANDIDATES, iters: int = 20) -> dict:
    from .sims import MultiGridSim
    out = {}
    for strat, pv in candidates:
        sim = MultiGridSim(cfg, workload=w, patch_verts=pv, strategy=strat)
        sim.engine.iterate(2)
        sim.engine.synchronize()
        sim.engine.kernel_timing(True)
        sim.engine.iterate(iters)
        ms, n = sim.engine.get_kernel_timing()
        sim.engine.kernel_timing(False)
        out[(strat, pv)] = ms * 1e3 / max(n, 1)
        sim.engine.close()
    return out


def pick(cfg: dict, w, use_cache: bool = True):
    """(strategy, patch_verts, {candidate: us per iteration})"""
    key = signature(cfg, w)
    cache = _load()
    if use_cache and key in cache:
        e = cache[key]
        return int(e["strategy"]), int(e["patch_verts"]), e["us_per_iteration"]
    times = time_candidates(cfg, w)
    best = min(times, key=times.get)
    table = {f"{'flat' if k[0] else 'patch'}{k[1] if not k[0] else ''}": round(v, 3) for k, v in times.items()}
    cache[key] = {"strategy": int(best[0]), "patch_verts": int(best[1]), "us_per_iteration": table}
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(cache, indent=1))
    except OSError:
        pass
    return int(best[0]), int(best[1]), table
