"""Launch-shape autotuner for the fused Jacobi projection.

The tunable is the patch size (owned vertices per patch = threads per
workgroup, 64/128/256): it trades workgroup count and occupancy against tet
replication across patch borders and halo loads.  Each candidate layout is
built for the actual mesh, a few iterations are timed with HIP events around
the projection launches, and the fastest is kept.  Results are cached per mesh
signature in ``rlmusclev2_amd/_tune/cache.json``.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

CACHE = Path(__file__).resolve().parent / "_tune" / "cache.json"
CANDIDATES = (64, 128, 256)


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
    for pv in candidates:
        sim = MultiGridSim(cfg, workload=w, patch_verts=pv)
        sim.engine.iterate(2)
        sim.engine.synchronize()
        sim.engine.kernel_timing(True)
        sim.engine.iterate(iters)
        ms, n = sim.engine.get_kernel_timing()
        sim.engine.kernel_timing(False)
        out[pv] = ms * 1e3 / max(n, 1)
        sim.engine.close()
    return out


def pick_patch_verts(cfg: dict, w, use_cache: bool = True) -> int:
    key = signature(cfg, w)
    cache = _load()
    if use_cache and key in cache:
        return int(cache[key]["patch_verts"])
    times = time_candidates(cfg, w)
    best = min(times, key=times.get)
    cache[key] = {"patch_verts": int(best), "us_per_iteration": {str(k): v for k, v in times.items()}}
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(cache, indent=1))
    except OSError:
        pass
    return int(best)
