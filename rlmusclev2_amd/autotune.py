"""Launch-shape autotuner for the fused Jacobi projection.

Candidates: the "patch" strategy (LDS-staged vertex table, one launch per
iteration) at 64/128/256 owned vertices per workgroup, which trades workgroup
count and occupancy against tet replication across patch borders; the "flat"
strategy (one thread per tet, then one per vertex; two launches per
iteration, no replication, every corner correction goes through memory); and
the "block" strategy (consecutive tets per workgroup, corner sums in LDS,
only vertices shared between blocks go through memory) at several block
sizes and thread counts, one launch per iteration or resident (every plain
iteration of a substep in one launch, one block per part keeping its
vertices in LDS; the iteration unroll is the whole substep).  Parts of
several blocks fit larger meshes but lose to the per-iteration launches on
them (fewer waves per SIMD), so only one block per part is weighed: it fits
meshes up to about 250 k tets, where it is the faster launch.  Each candidate layout is built for the actual mesh,
100 iterations are captured in one HIP graph and launched between two HIP
events (as they run inside a frame) from the same deformed state, and the fastest is kept (candidates
within 1 % of it are ties, which go to the largest block).  Results are cached per mesh signature in
``rlmusclev2_amd/_tune/cache.json``.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

from .capi import XpbdError

CACHE = Path(__file__).resolve().parent / "_tune" / "cache.json"
# (strategy, patch_verts, block_tets, block_threads, resident blocks per part)
CANDIDATES = ((1, 256, 0, 0, 0), (0, 64, 0, 0, 0), (0, 128, 0, 0, 0), (0, 256, 0, 0, 0),
              (2, 256, 128, 128, 0), (2, 256, 192, 192, 0), (2, 256, 256, 256, 0), (2, 256, 384, 384, 0),
              (2, 256, 512, 256, 0), (2, 256, 512, 512, 0), (2, 256, 768, 384, 0), (2, 256, 1024, 512, 0),
              (2, 256, 2048, 512, 0),
              (2, 256, 128, 128, 1), (2, 256, 192, 192, 1), (2, 256, 256, 256, 1))


def name(c) -> str:
    strat, pv, tb, bs, res = c
    if strat != 2:
        return "flat" if strat == 1 else f"patch{pv}"
    return f"block{tb}x{bs}" + (f"_resident{res}" if res else "")


def as_kwargs(c) -> dict:
    strat, pv, tb, bs, res = c
    return dict(strategy=strat, patch_verts=pv, block_tets=tb or 1024, block_threads=bs or 256,
                resident_blocks=res)


def signature(cfg: dict, w) -> str:
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(w.tets).tobytes()[:1 << 20])
    h.update(json.dumps([cfg.get("constraints"), w.nt, w.nv], sort_keys=True, default=str).encode())
    return h.hexdigest()[:16]


def choose(times: dict, tie: float = 0.01):
    """Fastest candidate; those within ``tie`` of it count as equal (run-to-run
    noise of the timing is about 1 %) and the tie goes to the largest block
    (fewer vertices shared between workgroups, less memory traffic), then to
    the faster one."""
    t0 = min(times.values())
    near = [c for c, t in times.items() if t <= t0 * (1.0 + tie)]
    return max(near, key=lambda c: (c[0] == 2, c[2] * max(c[3], 1), -times[c]))


def _load():
    try:
        return json.loads(CACHE.read_text())
    except Exception:
        return {}


def deformed_state(w, seed: int = 5) -> np.ndarray:
    """rest positions with every vertex moved by up to 5 % of the mean tet edge:
    a state whose closest rotations take a few polar steps, as in a frame (at
    rest every projection is cheap, which favours launch shapes with less
    arithmetic parallelism)"""
    p = np.asarray(w.pos0, np.float32).reshape(-1, 3)
    t = np.asarray(w.tets, np.int64).reshape(-1, 4)
    sample = t[:: max(1, len(t) // 20000)]
    edge = float(np.mean(np.linalg.norm(p[sample[:, 0]] - p[sample[:, 1]], axis=1)))
    rng = np.random.default_rng(seed)
    return (p + rng.uniform(-0.05 * edge, 0.05 * edge, p.shape)).astype(np.float32)


def time_candidates(cfg: dict, w, candidates=CANDIDATES, iters: int = 0) -> dict:
    """us per iteration of each candidate from the same deformed state, in launches
    of the config's iterations per substep (what a frame runs)"""
    from .sims import MultiGridSim
    n = iters or int(cfg.get("fine_iterations", 10))
    n = max(2, n + (n & 1))
    start = deformed_state(w)
    out = {}
    for c in candidates:
        try:
            sim = MultiGridSim(cfg, workload=w, **as_kwargs(c))
        except XpbdError:      # e.g. the layout needs more LDS than a workgroup can have
            continue
        sim.engine.set_state(pos=start, pprev=start)
        sim.engine.iterate(2)
        out[c] = sim.engine.time_iterations(n, max(1, 100 // n))
        sim.engine.close()
        print(f"autotune {name(c)}: {out[c]:.2f} us/iteration", file=sys.stderr, flush=True)
    return out


def pick(cfg: dict, w, use_cache: bool = True):
    """(MultiGridSim tuning kwargs, {candidate name: us per iteration})"""
    key = signature(cfg, w)
    cache = _load()
    if use_cache and key in cache and "tuning" in cache[key]:
        e = cache[key]
        return dict(e["tuning"]), e["us_per_iteration"]
    times = time_candidates(cfg, w)
    best = choose(times)
    table = {name(k): round(v, 3) for k, v in times.items()}
    cache[key] = {"tuning": as_kwargs(best), "us_per_iteration": table}
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        CACHE.write_text(json.dumps(cache, indent=1))
    except OSError:
        pass
    return as_kwargs(best), table
