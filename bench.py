#!/usr/bin/env python3
"""Benchmark of the XPBD muscle substep on one MI355X: the three parts of the
headline metric "ms/frame @535k tets; batched env-substeps/s; warm-start vs
plain @same residual".  Prints ONE JSON line.

* value: ms/frame of the humanV2-scale scene (534,882 tets: 40 spindle
  muscles + 4 fascia tubes, ARAP-SVD + volume, tendon-tip pins on 4 moving
  bones, 128^3 capsule SDFs; 10 substeps x 10 Jacobi iterations), one captured
  hipGraph per frame, ``--steps`` frames after ``--warmup``.  Inputs live in
  HBM before the timed region; the per-frame pin targets and bone transforms
  are uploaded inside it.
* "c4": 512 independent spindle instances (tets log-uniform in [2k, 40k],
  ARAP-SVD + volume + aniso_arap, a per-instance activation random walk
  uploaded every frame, one 64^3 capsule SDF each) in one context:
  instance-substeps/s, ms/frame, roofline.
* "c5": learned warm start against plain XPBD at the same residual, with
  self-collision DCD, on the humanV2-scale mesh and on a 64-instance uneven
  batch: network trained by the library's MFMA kernels on plain-solver frames,
  iterations to reach the plain 10-iteration residual, ms/frame of plain and
  warm frames with the audited per-frame fallback.
* "fascia": the humanV2-scale scene with its fascia tubes as the fascia
  sub-solver in ride mode (bound to the muscle surface, 20 Stable Neo-Hookean
  shell iterations per substep) and in squeeze mode (bone-driven shell, ARAP +
  volume with a shrunk rest shape, glue on the muscle surface), against the
  same muscles without it.
* "bone_mesh": the humanV2-scale frame with the exact-mesh bone collider (a
  capsule skeleton of 97,536 triangles, BVH refit per frame, closest-point
  push in every substep) in place of the grid SDFs, with the push, refit and
  penetration meter timed alone.

``--workload W`` runs one workload alone (the form the profiling commands use).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "ms/frame @535k tets; batched env-substeps/s; warm-start vs plain @same residual"

# HBM bytes per projection iteration measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per
# dispatch, scaled x2.0 and x1.0 by a plain-copy calibration), keyed by (tets, constraint types,
# strategy, block tets, block threads) for the per-iteration launches, with the blocks per part
# appended for the resident launch (its bytes per dispatch divided by the iterations it runs); the
# runs are summarised under profiles/
TRAFFIC = {
    (534882, ("tetarapsvd", "volume"), 2, 256, 256): 57575112.0,                    # humanV2-scale, plain
    (534882, ("tetarapsvd", "volume"), 2, 192, 192): 58638881.0,                    # the same, other shapes
    (534882, ("tetarapsvd", "volume"), 2, 384, 384): 56246416.0,
    (534882, ("tetarapsvd", "volume"), 2, 512, 512): 55199400.0,
    (534882, ("snh", "aniso_arap"), 2, 256, 256): 63921155.0,                       # humanV2-scale, SNH + aniso
    (7063998, ("tetarapsvd", "volume", "aniso_arap"), 2, 256, 256): 816899321.0,   # 512 instances
}

# VALU lane-instructions of one projection iteration: SQ_INSTS_VALU per wave x 64 lanes x waves of the tet
# kernel plus the shared-vertex kernel (rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU ..., 800 dispatches each;
# profiles/projection_valu_sq_counters.md), keyed like TRAFFIC
VALU_LANE_INSTS = {
    (534882, ("tetarapsvd", "volume"), 2, 256, 256): 642.0 * 64 * 8360 + 106.1 * 64 * 1488,
    (534882, ("tetarapsvd", "volume"), 2, 384, 384): 628.9 * 64 * 8358 + 103.4 * 64 * 1372,
    (534882, ("tetarapsvd", "volume"), 2, 192, 192): 653.1 * 64 * 8358 + 108.6 * 64 * 1568,
    (534882, ("tetarapsvd", "volume"), 2, 512, 512): 619.5 * 64 * 8360 + 101.0 * 64 * 1288,
    (7063998, ("tetarapsvd", "volume", "aniso_arap"), 2, 256, 256): 760.1 * 64 * 110376 + 99.8 * 64 * 18556,
}
# one FP32 VALU instruction per SIMD every 4 cycles for a 64-lane wave (16 lanes): 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz (packed FP32 is kept out of the build)
VALU_PEAK_LANE_INSTS = 256 * 4 * 16 * 2.4e9

WORKLOADS = {
    "humanv2": ("humanv2_scale.json", {}),
    "humanv2_20": ("humanv2_scale.json", {"fine_iterations": 20}),
    "snh_aniso": ("humanv2_snh_aniso.json", {}),
    "snh_aniso_voxel": ("humanv2_snh_aniso.json", {"mg_cage": "voxel"}),
    "cantilever": ("cantilever.json", {}),
    "batch": ("batch512.json", {}),
    "c5": ("humanv2_c5.json", {}),
    "c5_batch": ("batch64_c5.json", {}),
    "mesh_collider": ("humanv2_mesh_collider.json", {}),
    "fascia_ride": ("humanv2_fascia_ride.json", {}),
    "fascia_squeeze": ("humanv2_fascia_squeeze.json", {}),
}
# the block shape the autotuner picks on the 512-instance batch (tuned once per mesh; re-tuning it in
# every bench run would rebuild the 7 M-tet layout 13 times)
BATCH_TUNING = dict(strategy=2, patch_verts=256, block_tets=256, block_threads=256)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="all", choices=["all", *sorted(WORKLOADS)])
    ap.add_argument("--patch-verts", type=int, default=0, help="0 = autotune strategy and patch size")
    ap.add_argument("--strategy", type=int, default=-1, help="-1 = autotune, 0 = patch, 1 = flat, 2 = block")
    ap.add_argument("--block-tets", type=int, default=1024, help="block strategy: tets per workgroup")
    ap.add_argument("--block-threads", type=int, default=256, help="block strategy: threads per workgroup")
    ap.add_argument("--resident-blocks", type=int, default=0,
                    help="block strategy: 0 = one launch per iteration, R = resident mode with R blocks per "
                         "part, -1 = resident with the smallest R that fits")
    ap.add_argument("--c4-frames", type=int, default=10, help="timed frames of the 512-instance batch")
    ap.add_argument("--ws-train-frames", type=int, default=0, help="warm-start legs: training frames (0 = config)")
    ap.add_argument("--ws-eval-frames", type=int, default=0, help="warm-start legs: evaluation frames (0 = config)")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the warm-start legs")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every core this process may run on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_cfg(name):
    fn, over = WORKLOADS[name]
    cfg = json.loads((ROOT / "rlmusclev2_amd" / "configs" / fn).read_text())
    cfg.update(over)
    return cfg


class Drive:
    """Per-frame inputs that are not pins or bones: the activation ramp of
    RampActivationSource (src/muscle/util/activation.py:28-43) or, for batched
    instances, a seeded per-instance random walk clipped to [0, 1]."""

    def __init__(self, cfg, w, seed=21):
        import numpy as np
        self.np = np
        self.ramp = cfg.get("activation_ramp")
        self.walk = None
        has_fiber = any(c.get("name") == "fiber" and c.get("enable", True) for c in cfg.get("constraints", []))
        if has_fiber and not self.ramp and w.tet_instance is not None:
            from rlmusclev2_amd.warmstart import ActivationWalk
            self.walk = ActivationWalk(w.tet_instance, np.random.default_rng(seed))
        self.nt = w.nt

        self.groups_set = set()

    def __call__(self, sim, frame):
        if self.ramp:
            a = min(1.0, max(0.0, (frame - self.ramp[0]) / max(1, self.ramp[1] - self.ramp[0])))
            sim.set_fiber_activation(self.np.full(self.nt, a, "float32"))
        elif self.walk is not None:
            # one value per instance per frame, expanded to the tets on the device
            if id(sim) not in self.groups_set:
                sim.set_activation_groups(self.walk.g)
                self.groups_set.add(id(sim))
            sim.set_group_activation(self.walk.next_groups())


def run_frames(cfg, w, tuning, steps, warmup, label):
    """Build, warm up, time ``steps`` frames (graph replay each, host sync at both
    ends), then time the projection iterations alone on the library stream."""
    import numpy as np
    from rlmusclev2_amd import sims
    t0 = time.perf_counter()
    sim = sims.MultiGridSim(cfg, workload=w, **tuning)
    log(f"{label}: built in {time.perf_counter() - t0:.1f} s")
    st = sim.engine.stats()
    nsub, iters = sim.nsubsteps, sim.fine_iters
    drive = Drive(cfg, w)

    def frame(f):
        drive(sim, f)
        sim.step_frame(f)

    for f in range(1, warmup + 1):
        frame(f)
    sim.engine.synchronize()
    t0 = time.perf_counter()
    for f in range(warmup + 1, warmup + steps + 1):
        frame(f)
    sim.engine.synchronize()
    ms_frame = (time.perf_counter() - t0) * 1e3 / steps
    # the projection kernels alone: t_iters * t_reps iterations captured in one
    # HIP graph and launched once between two HIP events on the library stream
    # (the stream the kernels run on), from the state the timed frames left
    t_iters = iters if iters >= 2 and iters % 2 == 0 else 10
    t_reps = 20
    avg_us = sim.engine.time_iterations(t_iters, t_reps)
    strat = tuning["strategy"]
    res = int(tuning.get("resident_blocks", 0)) if strat == 2 else 0
    types = tuple(c["type"] for c in sim.scene.meta["constraint_configs"])
    key = (w.nt, types, strat, tuning.get("block_tets", 0) if strat == 2 else 0,
           tuning.get("block_threads", 0) if strat == 2 else 0)
    traffic = TRAFFIC.get(key + (st["resident_blocks"],) if res else key)
    valu = VALU_LANE_INSTS.get(key) if not res else None
    kernel = ("k_iterate (fused Jacobi projection)", "k_flat_tets + k_flat_verts (one Jacobi iteration)",
              "k_block_tets + k_block_shared (one Jacobi iteration)")[strat]
    if res:
        kernel = (f"k_resident (every plain iteration of a substep in one launch, {st['resident_parts']} parts "
                  f"of {st['resident_blocks']} blocks; per iteration)")
    alg = st["bytes_per_iter_alg"]
    achieved = alg / (avg_us * 1e-6) / 1e9
    rec = {
        "ms_per_frame": round(ms_frame, 4), "frames_timed": steps, "tets": w.nt, "verts": w.nv,
        "substeps": nsub, "iterations": iters, "constraints": list(types),
        "strategy": ("patch", "flat", "block")[strat], "tuning": tuning, "blocks": st["n_patches"],
        "tet_iterations_per_s": round(w.nt * iters * nsub / (ms_frame * 1e-3), 1),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kernel,
            "avg_iteration_us": round(avg_us, 3),
            "timing": "one HIP graph of the iteration kernels between two HIP events on the library stream",
            "iterations_timed": t_iters * t_reps, "alg_bytes_per_iteration": alg,
            "impl_bytes_per_iteration": st["bytes_per_iter_impl"],
            # what bounds it: VALU issue (FP32 lane-instructions per iteration from SQ counters)
            "valu": None if valu is None else {
                "lane_instructions_per_iteration": valu, "achieved_tlane_per_s": round(valu / avg_us * 1e-6, 2),
                "peak_tlane_per_s": round(VALU_PEAK_LANE_INSTS * 1e-12, 2),
                "frac": round(valu / (avg_us * 1e-6) / VALU_PEAK_LANE_INSTS, 4)},
        },
        "residual_rms": sim.residuals(),
        "finite": bool(np.all(np.isfinite(sim.positions_np()))),
    }
    if sim.scene.dcd is not None:
        rec["dcd"] = dict(sim.engine.dcd_stats(), surface_tris=int(sim.scene.dcd.tris.shape[0]))
    return sim, rec


def host_cpu():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail, os.cpu_count() or avail


def cpu_baseline(cfg, w, threads, budget_s=12.0):
    """The oracle (C restatement of the reference kernels, OpenMP) on host cores:
    whole substeps of frame 1 of the same scene until ~budget_s of wall time,
    scaled to ms/frame; then one substep on a single thread."""
    import oracle
    from rlmusclev2_amd import sims
    model, avail, total = host_cpu()

    def time_substeps(nthreads, budget, max_n):
        ref = sims.MultiGridSim(cfg, workload=w, engine_cls=lambda sc: oracle.OracleSim(sc, nthreads=nthreads))
        ref.load_frame_targets(1)
        ref.engine.apply_pintotarget()
        t0 = time.perf_counter()
        n = 0
        while True:
            ref.engine.native_substep(ref.fine_iters)
            n += 1
            if time.perf_counter() - t0 > budget or n >= max_n:
                break
        return (time.perf_counter() - t0) / n, n, ref.nsubsteps, ref.fine_iters

    s_all, n_all, nsub, iters = time_substeps(threads, budget_s, 20)
    s_one, n_one, _, _ = time_substeps(1, 0.0, 1)
    return {"value": round(s_all * nsub * 1e3, 3), "unit": "ms/frame", "cores": threads, "kind": "port",
            "sample": f"{n_all} substep(s) x {iters} Jacobi iterations of the same scene from frame 1 on {threads} "
                      f"OpenMP threads, {s_all * n_all:.1f} s wall, scaled to {nsub} substeps per frame",
            "one_thread_ms_per_frame": round(s_one * nsub * 1e3, 1),
            "one_thread_sample": f"{n_one} substep on 1 thread, scaled to {nsub} substeps",
            "cpu_model": model, "cores_available_to_process": avail, "cores_on_host": total,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def warm_start_leg(cfg, w, tuning, args):
    """Learned warm start on one mesh: collect plain-solver frames under
    randomized bone motion and activation walks, train the per-tet MLP with the
    HIP kernels, count the iterations each start needs to reach the plain
    10-iteration residual, then time plain frames against warm-started frames
    (audited per-frame fallback to plain) from the same trajectory start."""
    import numpy as np
    import torch
    from rlmusclev2_amd import sims, warmstart as ws
    wc = dict(cfg.get("warm_start", {}))
    if args.ws_train_frames:
        wc["train_frames"] = args.ws_train_frames
    if args.ws_eval_frames:
        wc["eval_frames"] = args.ws_eval_frames
    tuning = dict(tuning, strategy=tuning["strategy"] if tuning["strategy"] in (1, 2) else 1)
    sim = sims.MultiGridSim(cfg, workload=w, **tuning)
    t0 = time.perf_counter()
    ds = ws.collect(sim, n_frames=int(wc.get("train_frames", 2000)), seed=0,
                    samples_per_substep=int(wc.get("samples_per_substep", 4096)),
                    conv_iters=int(wc.get("conv_iters", 10)), lags=(1, 2), log=log)
    t1 = time.perf_counter()
    norm = ws.fit_normalizer(ds)
    net, hist = ws.train(ds, norm, epochs=int(wc.get("epochs", 3)), batch=16384, lr=2e-3, seed=0,
                         log=lambda h: log(f"train: {h}"))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n_samples = int(ds.X.shape[0])
    del ds
    sim.engine.close()
    ev = sims.MultiGridSim(cfg, workload=w, **tuning)
    ws.install(ev.engine, net, norm)
    log("evaluating iterations to the plain residual")
    cmp = ws.compare(ev, n_frames=int(wc.get("eval_frames", 5)), seed=11,
                     modes=(ws.MODE_PREV, ws.MODE_NET), conv_iters=int(wc.get("conv_iters", 10)))
    ev.engine.close()
    iters = {}
    for m, v in cmp.items():
        vv = np.array([11 if x is None else x for x in v])
        nm = {ws.MODE_PREV: "previous_lambda", ws.MODE_NET: "network", "exact": "exact_lambda_bound"}.get(m, str(m))
        iters[nm] = {"mean": round(float(vv.mean()), 3), "p90": float(np.percentile(vv, 90)),
                     "misses": int(sum(x is None for x in v)), "substeps": len(v)}
    nsub = cfg["nsubsteps"]
    R = int(wc.get("ref_every", 3))
    K = int(wc.get("k_iters", 4))
    nf = int(wc.get("timed_frames", 20))
    ws.randomize_motion(w, np.random.default_rng(13))
    runs = [("plain", None, None), ("network", K, True), ("previous_lambda", K, False)]
    tim = {}
    for label, k, use_net in runs:
        log(f"timing {label} frames (K={k}, reference solve every {R} substeps)")
        s2 = sims.MultiGridSim(cfg, workload=w, **tuning)
        ws.install(s2.engine, net, norm)
        drive = Drive(cfg, w)
        st = (ws.WarmStartStepper(s2, k, audit_every=int(wc.get("audit_every", 8)), tol=float(wc.get("tol", 0.0)),
                                  guard=None, ref_every=R, use_network=use_net) if k else None)

        def frame(f):
            drive(s2, f)
            st.step_frame(f) if st else s2.step_frame(f)

        for f in range(1, 3):
            frame(f)
        s2.engine.synchronize()
        if st:
            st.frames = st.fallbacks = st.audits = 0
        ta = time.perf_counter()
        for f in range(3, 3 + nf):
            frame(f)
        s2.engine.synchronize()
        ms = (time.perf_counter() - ta) * 1e3 / nf
        rec = {"ms_per_frame": round(ms, 4), "residual_end": s2.residuals()}
        if st:
            refs = (nsub + R - 1) // R
            rec.update(k_iters=k, ref_every=R, fallbacks=st.fallbacks, audits=st.audits,
                       iterations_per_frame=refs * st.plain_iters + (nsub - refs) * k)
        else:
            rec["iterations_per_frame"] = nsub * s2.fine_iters
        tim[label] = rec
        s2.engine.close()
    plain_ms = tim["plain"]["ms_per_frame"]
    for label in tim:
        if label != "plain":
            tim[label]["speedup_vs_plain"] = round(plain_ms / tim[label]["ms_per_frame"], 3)
    return {
        "workload": f"{w.name}: {w.nt} tets / {w.nv} verts" + (f", {w.info['n_inst']} instances"
                                                               if "n_inst" in w.info else ""),
        "train_frames": int(wc.get("train_frames", 2000)), "samples": n_samples,
        "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "last_epoch": hist[-1],
        "iterations_to_plain10_residual": iters, "timed_frames": nf,
        "fallback_tolerance": float(wc.get("tol", 0.0)), "audit_every": int(wc.get("audit_every", 8)),
        "frames": tim,
    }


def bone_mesh_leg(args, tuning, grid_ms):
    """The humanV2-scale frame with the exact-mesh bone collider in place of the
    grid SDFs (the reference's AnimatedBoneMeshCollider), with its pieces timed
    on the host around synchronised calls (launch included): one push of every
    vertex, the per-frame upload + refit, the penetration meter; with coherent
    queries (a vertex proven outside by its last query and the motion since is
    not queried again; same results) and with a query for every vertex."""
    import dataclasses
    from rlmusclev2_amd import sims, workloads
    cfg = load_cfg("mesh_collider")
    w = sims.build_workload(cfg)
    sim, rec = run_frames(cfg, w, tuning, args.steps, args.warmup, "mesh_collider")
    e = sim.engine
    bm = sim.scene.bone_mesh
    eps = float(cfg["demo"]["collision"].get("collide_eps", 1e-6))
    last = args.warmup + args.steps
    nrep = 20

    def per_call(fn, n=nrep):
        fn()
        e.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        e.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    _, n_in, max_d = e.bone_penetration(eps, want_depth=False)
    state = e.get_state()
    out = {}
    for mode, coh in (("coherent", bm.coherent), ("every_vertex", 0)):
        e.set_bone_mesh(dataclasses.replace(bm, verts=workloads.bone_mesh_points(w, last), coherent=coh))
        e.set_state(*state)
        meter_ms = per_call(lambda: e.bone_penetration(eps, want_depth=False), 10)
        push_ms = per_call(e.bone_mesh_push)
        e.set_state(*state)
        q0 = e.bone_mesh_queries()
        t0 = time.perf_counter()
        for f in range(last + 1, last + 4):
            sim.step_frame(f)
        e.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / 3
        out[mode] = {"ms_per_frame": round(ms, 4), "push_ms": round(push_ms, 4), "meter_ms": round(meter_ms, 4),
                     "queried_fraction_in_frames": round((e.bone_mesh_queries() - q0) / (3 * sim.nsubsteps * w.nv), 4)}
    pts = workloads.bone_mesh_points(w, last + 3)
    refit_ms = per_call(lambda: e.set_bone_mesh_points(pts))
    e.close()
    return {
        "workload": f"{w.name}: {w.nt} tets / {w.nv} verts against a {len(bm.tris)}-triangle skeleton mesh "
                    f"({bm.nbones} capsule bones), max_dist {bm.max_dist} m, contact_offset {bm.contact_offset} m",
        "ms_per_frame": rec["ms_per_frame"], "grid_sdf_ms_per_frame": grid_ms, **out,
        "refit_ms": round(refit_ms, 4), "inside_after_frames": n_in, "max_depth_after_frames": max_d,
        "timing": "frames: graph replays, host clock; push / meter / refit: host clock around synchronised calls",
        "finite": rec["finite"], "residual_rms": rec["residual_rms"],
    }


def fascia_leg(args, tuning):
    """The humanV2-scale scene with its fascia tubes as the fascia sub-solver in ride mode
    (the reference's humanV2 default, fascia.py:45-572): muscles 468 k tets, the shell
    67 k tets of Stable Neo-Hookean bound to the muscle surface, 20 shell iterations per
    substep; against the same muscles without the shell."""
    from rlmusclev2_amd import sims
    cfg = load_cfg("fascia_ride")
    w = sims.build_workload(cfg)
    sim, rec = run_frames(cfg, w, tuning, args.steps, args.warmup, "fascia_ride")
    face, _, _, nb = sim.engine.fascia_bind()
    fs = sim.fascia
    sim.engine.close()
    fs.engine.close()
    sim0, rec0 = run_frames(dict(cfg, use_fascia=False), w, tuning, args.steps, args.warmup, "fascia_ride_off")
    sim0.engine.close()
    cfg_s = load_cfg("fascia_squeeze")
    sim_s, rec_s = run_frames(cfg_s, w, tuning, args.steps, args.warmup, "fascia_squeeze")
    fss = sim_s.fascia
    sim_s.engine.close()
    fss.engine.close()
    return {
        "workload": f"{w.name} muscles {w.nt} tets / {w.nv} verts; fascia shell {fs.NT} tets / {fs.NV} verts, "
                    f"{nb} bound to {len(fs.surf_tris)} muscle surface triangles (bind_max_dist "
                    f"{fs.bind_max_dist} m), {fs.iters} shell iterations per substep",
        "ms_per_frame": rec["ms_per_frame"], "ms_per_frame_without_fascia": rec0["ms_per_frame"],
        "fascia_ms_per_frame": round(rec["ms_per_frame"] - rec0["ms_per_frame"], 4),
        "fascia_tet_iterations_per_s": round(fs.NT * fs.iters * sim.nsubsteps /
                                             max(1e-9, (rec["ms_per_frame"] - rec0["ms_per_frame"]) * 1e-3), 1),
        "bound": nb, "finite": rec["finite"],
        "squeeze": {"ms_per_frame": rec_s["ms_per_frame"],
                    "fascia_ms_per_frame": round(rec_s["ms_per_frame"] - rec0["ms_per_frame"], 4),
                    "anchors": int(len(fss.anchor)), "iterations": fss.iters, "finite": rec_s["finite"]},
    }


def one_workload(args, name):
    """A single workload (its own line; the profiling commands use this form)."""
    from rlmusclev2_amd import autotune, sims
    cfg = load_cfg(name)
    w = sims.build_workload(cfg)
    tune_table = None
    if args.patch_verts or args.strategy >= 0:
        tuning = dict(strategy=max(args.strategy, 0), patch_verts=args.patch_verts or 256,
                      block_tets=args.block_tets, block_threads=args.block_threads,
                      resident_blocks=args.resident_blocks)
    elif name == "batch":
        tuning = dict(BATCH_TUNING)
    else:
        tuning, tune_table = autotune.pick(cfg, w)
    if name == "mesh_collider":
        del w
        return {"metric": "ms/frame (humanV2-scale, exact-mesh bone collider)", "unit": "ms/frame",
                "bone_mesh": bone_mesh_leg(args, tuning, None)}
    if name == "fascia_ride":
        del w
        return {"metric": "ms/frame (humanV2-scale, fascia shell in ride mode)", "unit": "ms/frame",
                "fascia": fascia_leg(args, tuning)}
    sim, rec = run_frames(cfg, w, tuning, args.steps, args.warmup, name)
    sim.engine.close()
    rec["autotune_us_per_iteration"] = tune_table
    n_inst = int(w.info.get("n_inst", 1))
    rec["env_substeps_per_s"] = round(n_inst * rec["substeps"] / (rec["ms_per_frame"] * 1e-3), 1)
    out = {"metric": f"ms/frame ({name})", "value": rec["ms_per_frame"], "unit": "ms/frame", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["ms_per_frame"],
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded generator, no assets)",
           "config": {"workload": f"{w.name}: {w.nt} tets / {w.nv} verts"}, **rec}
    if "warm_start" in cfg and not args.no_warm_start:
        out["warm_start"] = warm_start_leg(cfg, w, tuning, args)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; multi-GPU sharding is a design note only "
                                   "(instances or mesh partitions per GPU), --gpus must be 1"}))
        return 2
    import torch  # device memory is owned by the HIP library; torch runs the copy calibration and the MLP data

    if args.calibrate_copy_mb:
        n = args.calibrate_copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.calibrate_copy_mb, "bytes_read": 4 * n, "bytes_written": 4 * n}))
        return 0
    if args.workload != "all":
        print(json.dumps(one_workload(args, args.workload)))
        return 0

    from rlmusclev2_amd import autotune, sims
    # ---- headline: humanV2-scale plain solve ------------------------------
    cfg = load_cfg("humanv2")
    w = sims.build_workload(cfg)
    if args.patch_verts or args.strategy >= 0:
        tuning = dict(strategy=max(args.strategy, 0), patch_verts=args.patch_verts or 256,
                      block_tets=args.block_tets, block_threads=args.block_threads,
                      resident_blocks=args.resident_blocks)
        tune_table = None
    else:
        tuning, tune_table = autotune.pick(cfg, w)
    sim, rec = run_frames(cfg, w, tuning, args.steps, args.warmup, "humanv2")
    sim.engine.close()
    out = {
        "metric": METRIC, "value": rec["ms_per_frame"], "unit": "ms/frame", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rec["ms_per_frame"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator, no assets)",
        "config": {
            "workload": f"{w.name}: {w.nt} tets / {w.nv} verts, {rec['substeps']} substeps x {rec['iterations']} "
                        f"Jacobi iterations, constraints {rec['constraints']}, "
                        f"{len(w.pin_vids) if w.pin_vids is not None else 0} pinned verts, {len(w.bones)} bone SDFs",
            "tets": w.nt, "verts": w.nv, "substeps": rec["substeps"], "iterations": rec["iterations"],
            "strategy": rec["strategy"], "tuning": tuning, "blocks": rec["blocks"],
            "autotune_us_per_iteration": tune_table,
        },
        "tet_iterations_per_s": rec["tet_iterations_per_s"],
        "roofline": rec["roofline"],
        "residual_rms": rec["residual_rms"],
        "finite": rec["finite"],
    }
    # ---- c4: 512 instances in one context --------------------------------
    cfg4 = load_cfg("batch")
    w4 = sims.build_workload(cfg4)
    sim4, rec4 = run_frames(cfg4, w4, dict(BATCH_TUNING), args.c4_frames, 2, "batch512")
    sim4.engine.close()
    n_inst = int(w4.info["n_inst"])
    out["c4"] = {
        "workload": f"{n_inst} spindle instances, {w4.nt} tets / {w4.nv} verts, tets per instance log-uniform in "
                    f"[{cfg4['synthetic']['min_tets']}, {cfg4['synthetic']['max_tets']}], per-instance activation "
                    f"walk uploaded per frame, one {cfg4['synthetic']['sdf_res']}^3 capsule SDF per instance",
        "instances": n_inst,
        "env_substeps_per_s": round(n_inst * rec4["substeps"] / (rec4["ms_per_frame"] * 1e-3), 1),
        **rec4,
    }
    del w4
    # ---- exact-mesh bone collider on the headline mesh -------------------
    out["bone_mesh"] = bone_mesh_leg(args, tuning, rec["ms_per_frame"])
    # ---- the fascia sub-solver (ride) on the same scene --------------------
    out["fascia"] = fascia_leg(args, tuning)
    # ---- c5: warm start vs plain at the same residual, with DCD ----------
    if not args.no_warm_start:
        out["c5"] = {}
        for key, name in (("humanv2", "c5"), ("batch64", "c5_batch")):
            cfg5 = load_cfg(name)
            w5 = sims.build_workload(cfg5)
            t5 = dict(BATCH_TUNING) if name == "c5_batch" else tuning
            log(f"warm-start leg {name}")
            out["c5"][key] = warm_start_leg(cfg5, w5, t5, args)
    # ---- CPU baseline: the oracle on the host cores ----------------------
    if not args.no_cpu_baseline:
        # every core this process may run on, within the OpenMP share the host grants it
        avail = host_cpu()[1]
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = args.cpu_threads or (min(avail, omp) if omp > 0 else avail)
        out["cpu_baseline"] = cpu_baseline(cfg, w, threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
