"""Learned warm start for the Jacobi XPBD substep (no reference counterpart).

The reference GPU path is "plain zero-lambda XPBD": every Jacobi iteration
starts from lambda = 0 (solve_kernels.py:271-281, clear() each iteration at
mg.py:1348).  Here the first iteration of a substep applies a multiplier
estimate lambda0 as an impulse along the constraint gradients and the
remaining iterations are plain zero-lambda iterations from that start (a
carried lambda would converge towards the compliant solution, whose residual
alpha * lambda is above the plain solver's).  A small per-tet MLP predicts lambda0 as a
per-constraint-type gain on the previous substep's final multipliers,
``lambda0 = lambda_prev * clamp(1 + y, 0, 2)``, from local state at the start
of the substep (deformation gradient, fibre data, activation, the previous
multipliers, rest volume, compliances).  A gain keeps the previous
multipliers' spatial coherence: independent per-tet errors in an absolute
prediction inject high-frequency strain that Jacobi removes slowly.

Training data comes from plain-solver trajectories with randomized activations
and bone motions.  The target is the multipliers that a lambda-carrying solve
of ``conv_iters`` iterations (10, the plain iteration count) accumulates from
each substep start: applied as the first-iteration impulse, those multipliers
reach the plain 10-iteration residual in about 4 iterations (measured with the
exact multipliers of the same substep); longer solves accumulate multipliers
whose one-shot impulse overshoots.  Plain XPBD is both the yardstick (its
10-iteration residual is the target) and the per-frame fallback.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import workloads
from .mlp import NIN, NOUT, Mlp

MODE_PLAIN, MODE_CARRY, MODE_NET, MODE_PREV, MODE_SCHED = 0, 1, 2, 3, 4


def _dptr(t):
    return C.c_void_p(t.data_ptr())


def randomize_motion(w, rng, amp_deg=(4.0, 12.0), freq=(0.3, 0.8)):
    """New bone phases/amplitudes/frequencies (|dtheta/dt| <= 30 deg/s kept by clipping)."""
    for bn in w.bones:
        f = float(rng.uniform(*freq))
        amp = np.deg2rad(rng.uniform(*amp_deg))
        amp = min(amp, np.deg2rad(30.0) / (2 * np.pi * f))
        bn.update(amp=amp, freq=f, phase=float(rng.uniform(0, 2 * np.pi)))


class ActivationWalk:
    """Per-muscle (or per-instance) activation random walk clipped to [0, 1]."""

    def __init__(self, tet_group, rng, step=0.08):
        self.g = np.asarray(tet_group)
        self.n = int(self.g.max()) + 1 if self.g.size else 1
        self.rng = rng
        self.step = step
        self.a = rng.uniform(0, 1, self.n).astype(np.float32)

    def next(self):
        return self.next_groups()[self.g]

    def next_groups(self):
        """the next step of the walk, one value per group"""
        self.a = np.clip(self.a + self.rng.normal(0, self.step, self.n), 0, 1).astype(np.float32)
        return self.a


def tet_groups(w):
    if w.tet_instance is not None:
        return w.tet_instance
    mid = w.info.get("muscle_id")
    if mid is None:
        return np.zeros(w.nt, np.int32)
    return mid[w.tets[:, 0]]


@dataclass
class Dataset:
    X: object            # torch (N, 24) raw features
    LP: object           # torch (N, 4) previous-substep multipliers
    LT: object           # torch (N, 4) target multipliers of this substep
    meta: dict = field(default_factory=dict)


def collect(sim, n_frames: int, seed: int = 0, samples_per_substep: int = 8192, conv_iters: int = 10,
            substep_stride: int = 2, lags=(1,), log=None) -> Dataset:
    """Plain-solver trajectory with randomized activations and bone motion.

    At every substep, after integrate, a lambda-carrying solve of
    ``conv_iters`` iterations from lambda = 0 gives the converged multipliers
    (the state is restored afterwards, so the trajectory stays plain).  Every
    ``substep_stride``-th substep records, for each lag in ``lags``, (features
    with the converged multipliers of that many substeps back as the base,
    this substep's converged multipliers): a schedule with a reference solve
    every third substep applies a base one or two substeps old."""
    import torch
    e = sim.engine
    w = sim.workload
    rng = np.random.default_rng(seed)
    randomize_motion(w, rng)
    walk = ActivationWalk(tet_groups(w), rng)
    ne = e.stats()["n_elems"]
    p0 = sim.scene.pos0
    e.set_state(p0, p0, np.zeros_like(p0))
    feat = torch.empty((ne, NIN), dtype=torch.float32, device="cuda")
    lam = torch.zeros((ne, 4), dtype=torch.float32, device="cuda")
    lam_prev = torch.zeros((ne, 4), dtype=torch.float32, device="cuda")
    lam_prev2 = torch.zeros_like(lam_prev)
    feat2 = torch.empty_like(feat) if 2 in lags else None
    gen = torch.Generator(device="cuda").manual_seed(seed)
    Xs, Ps, Ts = [], [], []
    has_fib = sim.scene.v_fiber_dir is not None
    t0 = time.perf_counter()
    sub = 0
    for f in range(1, n_frames + 1):
        if has_fib:
            e.set_fiber_activation(walk.next())
        sim.load_frame_targets(f)
        e.apply_pintotarget()
        for s in range(sim.nsubsteps):
            e.set_lambda_mode(MODE_PLAIN)
            e.integrate()
            rec = sub % substep_stride == 0
            if rec and feat2 is not None:
                e.set_lambda(lam_prev2)
                e.features(feat2, normalized=False)
            e.set_lambda(lam_prev)
            if rec:
                e.features(feat, normalized=False)
            e.snapshot(True)
            e.set_lambda_mode(MODE_CARRY)
            e.warm_start()
            e.iterate(conv_iters)
            e.get_lambda(lam)
            e.snapshot(False)
            e.set_lambda_mode(MODE_PLAIN)
            if rec:
                idx = torch.randint(0, ne, (min(samples_per_substep, ne),), device="cuda", generator=gen)
                if 1 in lags:
                    Xs.append(feat[idx].clone())
                    Ps.append(lam_prev[idx].clone())
                    Ts.append(lam[idx].clone())
                if feat2 is not None:
                    Xs.append(feat2[idx].clone())
                    Ps.append(lam_prev2[idx].clone())
                    Ts.append(lam[idx].clone())
            lam_prev2.copy_(lam_prev)
            lam_prev.copy_(lam)
            e.iterate(sim.fine_iters)
            e.collide()
            e.update_velocities()
            sub += 1
        if log and f % 100 == 0:
            log(f"collect: frame {f}/{n_frames}, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    return Dataset(torch.cat(Xs), torch.cat(Ps), torch.cat(Ts),
                   dict(frames=n_frames, samples=int(sum(x.shape[0] for x in Xs)),
                        seconds=time.perf_counter() - t0, conv_iters=conv_iters))


@dataclass
class Normalizer:
    mean: np.ndarray
    invstd: np.ndarray
    oscale: np.ndarray


def fit_normalizer(ds: Dataset) -> Normalizer:
    """Feature standardisation; oscale = 1 (the network output is the gain - 1)."""
    X = ds.X.double()
    mean = X.mean(0)
    std = X.std(0)
    invstd = torch_where(std > 1e-6 * (mean.abs() + 1e-30), 1.0 / std.clamp_min(1e-30), 0.0)
    return Normalizer(mean.float().cpu().numpy(), invstd.float().cpu().numpy(), np.ones(NOUT, np.float32))


def torch_where(cond, a, b):
    import torch
    return torch.where(cond, a, torch.full_like(a, b))


def _loss_and_grad(yp, lp, lt, w):
    """loss = mean_n sum_c w_c (lp (1 + y) - lt)^2 over the clamp's open range;
    returns (loss, dloss/dy)."""
    g = 1.0 + yp
    inside = ((g > 0.0) & (g < 2.0)).float()
    r = lp * g.clamp(0.0, 2.0) - lt
    n = yp.shape[0]
    loss = float((w * r * r).sum() / n)
    dy = (2.0 / n) * w * r * lp * inside
    return loss, dy


def train(ds: Dataset, norm: Normalizer, epochs: int = 30, batch: int = 16384, lr: float = 2e-3,
          seed: int = 0, log=None) -> tuple[Mlp, list]:
    """Weighted squared error of the gained multipliers; forward and backward
    are the HIP MFMA kernels, Adam runs in torch."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    mean = torch.tensor(norm.mean, device="cuda")
    inv = torch.tensor(norm.invstd, device="cuda")
    X = ((ds.X - mean) * inv).contiguous()
    LP, LT = ds.LP.contiguous(), ds.LT.contiguous()
    # per-type weights: 1 / mean square target (types with no samples weigh 0)
    ms = (LT * LT).mean(0)
    w = torch.where(ms > 0, 1.0 / ms.clamp_min(1e-30), torch.zeros_like(ms))
    n = X.shape[0]
    nval = max(1024, n // 20)
    perm = torch.randperm(n, generator=g, device="cuda")
    val, tr = perm[:nval], perm[nval:]
    net = Mlp(seed=seed)
    with torch.no_grad():   # start from gain 1 (lambda0 = lambda_prev)
        net.p["W3"].mul_(0.01)
        net.p["b3"].zero_()
        net._refresh()
    hist = []
    steps_per_epoch = max(1, tr.numel() // batch)
    total = epochs * steps_per_epoch
    step = 0
    for ep in range(epochs):
        order = tr[torch.randperm(tr.numel(), generator=g, device="cuda")]
        for b in range(steps_per_epoch):
            idx = order[b * batch:(b + 1) * batch]
            xb = X[idx].contiguous()
            yp, h1, h2 = net.forward(xb, keep=True)
            _, dy = _loss_and_grad(yp, LP[idx], LT[idx], w)
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * step / total))
            net.adam_step(net.backward(xb, h1, h2, dy.contiguous()), lr=lr_t)
            step += 1
        yv = net.forward(X[val].contiguous())
        vl, _ = _loss_and_grad(yv, LP[val], LT[val], w)
        v_prev, _ = _loss_and_grad(torch.zeros_like(yv), LP[val], LT[val], w)
        v_zero = float((w * LT[val] * LT[val]).sum() / nval)
        hist.append({"epoch": ep, "val_loss": vl, "prev_loss": v_prev, "zero_loss": v_zero})
        if log:
            log(hist[-1])
    return net, hist


def install(engine, net: Mlp, norm: Normalizer):
    p = net.numpy_params()
    arr = {k: np.ascontiguousarray(v, np.float32) for k, v in p.items()}
    mean = np.ascontiguousarray(norm.mean, np.float32)
    inv = np.ascontiguousarray(norm.invstd, np.float32)
    osc = np.ascontiguousarray(norm.oscale, np.float32)
    engine.set_mlp(arr, mean, inv, osc)


def save(path, net: Mlp, norm: Normalizer, meta: dict | None = None):
    p = net.numpy_params()
    np.savez(path, **{k: v.astype(np.float32) for k, v in p.items()}, feat_mean=norm.mean,
             feat_invstd=norm.invstd, out_scale=norm.oscale, meta=np.array(str(meta or {})))


def load(path):
    z = np.load(path, allow_pickle=False)
    p = {k: z[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
    return Mlp(p), Normalizer(z["feat_mean"], z["feat_invstd"], z["out_scale"])


def _res(e):
    r = e.residuals()
    return r


def iterations_to_reach(sim, mode: int, target: dict, kmax: int = 10):
    """From the current (integrated) state: run ``mode`` one iteration at a time
    and return the first K whose every per-type residual is <= target, else None.
    State is restored afterwards."""
    e = sim.engine
    e.snapshot(True)
    e.set_lambda_mode(mode)
    e.warm_start()
    hit = None
    for k in range(1, kmax + 1):
        e.iterate(1)
        r = _res(e)
        if all(r[t] <= target[t] for t in target):   # False for NaN
            hit = k
            break
    e.snapshot(False)
    e.set_lambda_mode(MODE_PLAIN)
    return hit


def compare(sim, n_frames: int, seed: int = 1, plain_iters: int = 10, modes=(MODE_CARRY, MODE_PREV, MODE_NET),
            conv_iters: int = 10, exact: bool = True, log=None):
    """Per substep of a fresh randomized plain trajectory: the plain
    ``plain_iters``-iteration residual is the target; record the iterations
    each mode needs to reach it (None = not within ``plain_iters``).  Every
    mode starts from the previous substep's converged multipliers.  With
    ``exact``, key "exact" is the bound a perfect predictor would give
    (lambda0 = this substep's own converged multipliers)."""
    import torch
    e = sim.engine
    w = sim.workload
    rng = np.random.default_rng(seed)
    randomize_motion(w, rng)
    walk = ActivationWalk(tet_groups(w), rng)
    has_fib = sim.scene.v_fiber_dir is not None
    ne = e.stats()["n_elems"]
    lam_prev = torch.zeros((ne, 4), dtype=torch.float32, device="cuda")
    lam = torch.zeros_like(lam_prev)
    out = {m: [] for m in modes}
    if exact:
        out["exact"] = []
    p0 = sim.scene.pos0
    e.set_state(p0, p0, np.zeros_like(p0))
    for f in range(1, n_frames + 1):
        if has_fib:
            e.set_fiber_activation(walk.next())
        sim.load_frame_targets(f)
        e.apply_pintotarget()
        for s in range(sim.nsubsteps):
            e.set_lambda_mode(MODE_PLAIN)
            e.integrate()
            e.snapshot(True)
            e.iterate(plain_iters)
            target = _res(e)
            e.snapshot(False)
            e.snapshot(True)
            e.set_lambda_mode(MODE_CARRY)
            e.warm_start()
            e.iterate(conv_iters)
            e.get_lambda(lam)
            e.snapshot(False)
            for m in modes:
                e.set_lambda(lam_prev)
                out[m].append(iterations_to_reach(sim, m, target, kmax=plain_iters))
            if exact:
                e.set_lambda(lam)
                out["exact"].append(iterations_to_reach(sim, MODE_PREV, target, kmax=plain_iters))
            lam_prev.copy_(lam)
            e.set_lambda_mode(MODE_PLAIN)
            e.iterate(plain_iters)
            e.collide()
            e.update_velocities()
            if log:
                log(f"compare: frame {f} substep {s}: " + ", ".join(f"{m}={out[m][-1]}" for m in out))
    return out


class WarmStartStepper:
    """Frames with the learned warm start and the per-frame fallback to plain
    zero-lambda XPBD.

    A warm frame runs in mode 4: its reference substeps (the first one, or
    every ``ref_every``-th) are solves with lambda carried from zero and
    ``plain_iters`` iterations whose multipliers become the base; every other
    substep applies the base as an impulse (scaled by the network when
    ``use_network``) and runs ``k_iters`` iterations in all.  An audit
    frame (every ``audit_every``-th frame; 1 = every frame, the exact rule)
    also runs the plain ``plain_iters`` frame from the same start state and
    compares the per-type residuals at the end of the frame: if any type of
    the warm frame is worse by more than ``tol`` (or not finite), the frame is
    re-run plain and the warm start stays off until the next audit that it
    passes.  Between audits a warm frame still falls back at once when its
    residual is not finite or exceeds the last audited plain residual by more
    than ``guard``; with ``guard=None`` frames between audits are not checked
    (no residual readback, so the host does not wait for the GPU)."""

    def __init__(self, sim, k_iters: int, plain_iters: int | None = None, audit_every: int = 1,
                 tol: float = 0.0, guard: float | None = 1.0, use_graph: bool = True, ref_every: int = 0,
                 use_network: bool = True):
        self.sim = sim
        self.k = int(k_iters)
        self.plain_iters = int(plain_iters or sim.fine_iters)
        self.ref_every = int(ref_every)
        sim.engine.set_warm_schedule(self.plain_iters, self.ref_every, use_network)
        self.audit_every = max(1, int(audit_every))
        self.tol = float(tol)
        self.guard = None if guard is None else float(guard)
        self.use_graph = use_graph
        self.ref = None
        self.warm_ok = True
        self.frames = 0
        self.fallbacks = 0
        self.audits = 0

    def _frame(self, mode: int, iters: int, residual: bool = True):
        e = self.sim.engine
        e.set_lambda_mode(mode)
        e.native_frame(self.sim.nsubsteps, iters, self.use_graph)
        e.set_lambda_mode(MODE_PLAIN)
        return e.residuals() if residual else None

    @staticmethod
    def _worse(r, ref, tol):
        return not all(r[t] <= ref[t] * (1.0 + tol) for t in ref)   # NaN counts as worse

    def step_frame(self, frame: int) -> dict:
        sim, e = self.sim, self.sim.engine
        sim.load_frame_targets(frame)
        e.apply_pintotarget()
        audited = self.frames % self.audit_every == 0
        r = None
        if audited:
            e.snapshot(True)
            self.ref = self._frame(MODE_PLAIN, self.plain_iters)
            self.audits += 1
            e.snapshot(False)
            r = self._frame(MODE_SCHED, self.k)
            self.warm_ok = not self._worse(r, self.ref, self.tol)
            fallback = not self.warm_ok
        elif self.warm_ok and self.guard is None:
            self._frame(MODE_SCHED, self.k, residual=False)
            fallback = False
        elif self.warm_ok:
            e.snapshot(True)
            r = self._frame(MODE_SCHED, self.k)
            fallback = self._worse(r, self.ref, self.guard)
        else:
            fallback = True
        if fallback:
            if r is not None:
                e.snapshot(False)          # back to the start of the frame
            self._frame(MODE_PLAIN, self.plain_iters)
            self.fallbacks += 1
        self.frames += 1
        return {"fallback": fallback, "audited": audited, "residual": r, "plain": dict(self.ref or {})}
