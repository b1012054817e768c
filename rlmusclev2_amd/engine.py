"""XpbdEngine on the HIP substep library.

Same primitive names and semantics as the reference ``XpbdEngine``
(src/muscle/engine.py:785-1134): ``integrate``, ``clear``,
``solve_constraints``, ``apply_dP``, ``apply_pintotarget``, ``solve_extras``,
``post_smooth``, ``update_velocities``, ``run_program``.  Each maps to one C-ABI
call (include/xpbd.h).  The hot path is ``substep`` / ``step_frame``: the whole
substep (integrate + pin, ``iters`` fused Jacobi iterations, bone push-out,
velocity update) runs as native launches, and a frame replays one captured
hipGraph.

Only Jacobi mode exists on the GPU (the reference also forces Jacobi there,
musclesim.py:72-73); ``use_jacobi=False`` raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import XpbdError, f32, i32, ptr
from .scene import Scene


class XpbdEngine:
    """Device state + phase primitives for one mesh (or one packed batch)."""

    def __init__(self, scene: Scene, device: int = 0, patch_verts: int = 256, strategy: int = 0,
                 block_tets: int = 1024, block_threads: int = 256, resident_blocks: int = 0):
        self.scene = scene
        self.ctx = capi.Context(device)
        self.n_verts = scene.nv
        self.NV = scene.nv
        self.NT = scene.nt
        self.dt = float(scene.params.dt)
        self.use_jacobi = True
        self.collision_solvers = []        # extra Python post-phase hooks (break graph capture)
        self.collision_solvers_iter = []   # per-iteration Python hooks
        self.constraint_solvers = []       # MuscleSim extras
        self._keep = {}
        self.strategy = int(strategy)
        self.block_tets = int(block_tets)
        self.block_threads = int(block_threads)
        self.resident_blocks = int(resident_blocks)
        self._upload(scene, patch_verts)

    # -- setup -----------------------------------------------------------
    def _upload(self, s: Scene, patch_verts: int):
        c = self.ctx
        k = self._keep
        k["pos0"] = f32(s.pos0, (-1, 3))
        k["tets"] = i32(s.tets, (-1, 4))
        c.call("xpbd_set_mesh", s.nv, ptr(k["pos0"]), s.nt, ptr(k["tets"]), C.c_float(s.density))
        if s.mass is not None:
            k["mass"] = f32(s.mass)
            c.call("xpbd_set_mass", ptr(k["mass"]))
        if s.stopped is not None:
            k["stopped"] = i32(s.stopped)
            c.call("xpbd_set_stopped", ptr(k["stopped"]))
        if s.v_fiber_dir is not None or s.tendonmask is not None:
            k["vf"] = None if s.v_fiber_dir is None else f32(s.v_fiber_dir, (-1, 3))
            k["tm"] = None if s.tendonmask is None else f32(s.tendonmask)
            c.call("xpbd_set_fiber", ptr(k["vf"]), ptr(k["tm"]))
        t = s.cons.contiguous()
        c.call("xpbd_upload_constraints", t.n, ptr(t.type), ptr(t.pts), ptr(t.stiffness),
               ptr(t.compressionstiffness), ptr(t.dampingratio), ptr(t.tetid),
               ptr(t.restlength), ptr(t.restvector), ptr(t.restdir))
        self.set_params(s.params)
        if s.pin_vids is not None:
            k["pin_vids"] = i32(s.pin_vids)
            k["pin_active"] = i32(s.pin_active if s.pin_active is not None
                                  else np.ones(len(s.pin_vids), np.int32))
            c.call("xpbd_set_pins", len(k["pin_vids"]), ptr(k["pin_vids"]), ptr(k["pin_active"]))
        if s.sdf is not None:
            sd = s.sdf
            k["grids"] = f32(sd.grids)
            k["origins"] = f32(sd.origins, (-1, 3))
            k["spacing"] = f32(sd.spacing)
            k["coll_w"] = None if sd.coll_w is None else f32(sd.coll_w)
            c.call("xpbd_set_sdf", sd.nbones, sd.res, ptr(k["grids"]), ptr(k["origins"]),
                   ptr(k["spacing"]), C.c_float(sd.contact_offset), C.c_float(sd.max_push),
                   ptr(k["coll_w"]))
            if sd.vertex_group is not None:
                k["goff"] = i32(sd.group_bone_off)
                k["vgroup"] = i32(sd.vertex_group)
                c.call("xpbd_set_sdf_groups", len(k["goff"]) - 1, ptr(k["goff"]), ptr(k["vgroup"]))
        if s.dcd is not None:
            self._set_dcd(s.dcd)
        if s.rest_scale != 1.0:
            c.call("xpbd_scale_rest_shape", C.c_float(s.rest_scale))
        if s.bone_mesh is not None:
            self.set_bone_mesh(s.bone_mesh)
        tun = capi.XpbdTuning()
        tun.patch_verts = int(patch_verts)
        tun.strategy = self.strategy
        tun.block_tets = self.block_tets
        tun.block_threads = self.block_threads
        tun.resident_blocks = self.resident_blocks
        c.call("xpbd_finalize", C.byref(tun))
        if s.pin_vids is not None:
            tgt = f32(s.pos0, (-1, 3))[k["pin_vids"]]
            self.set_pin_targets(tgt)

    # exact-mesh bone collider (AnimatedBoneMeshCollider, humanV2.py:319-418) -----
    def set_bone_mesh(self, bm) -> None:
        """Install (or with None remove) the bone-mesh collider; allowed after finalize."""
        if bm is None:
            self.ctx.call("xpbd_set_bone_mesh", None)
            return
        k = self._keep
        k["bm_verts"] = f32(bm.verts, (-1, 3))
        k["bm_tris"] = i32(bm.tris, (-1, 3))
        k["bm_toff"] = None if bm.tri_off is None else i32(bm.tri_off)
        k["bm_goff"] = None if bm.group_bone_off is None else i32(bm.group_bone_off)
        k["bm_vgroup"] = None if bm.vertex_group is None else i32(bm.vertex_group)
        k["bm_cw"] = None if bm.coll_w is None else f32(bm.coll_w)
        k["bm_mw"] = None if bm.meter_w is None else f32(bm.meter_w)
        d = capi.XpbdBoneMeshDesc()
        d.nverts = k["bm_verts"].shape[0]
        d.verts = ptr(k["bm_verts"])
        d.ntris = k["bm_tris"].shape[0]
        d.tris = ptr(k["bm_tris"])
        d.nbones = bm.nbones
        d.tri_off = ptr(k["bm_toff"])
        d.ngroups = 0 if k["bm_goff"] is None else len(k["bm_goff"]) - 1
        d.goff = ptr(k["bm_goff"])
        d.vgroup = ptr(k["bm_vgroup"])
        d.coll_w = ptr(k["bm_cw"])
        d.meter_w = ptr(k["bm_mw"])
        d.max_dist = float(bm.max_dist)
        d.contact_offset = float(bm.contact_offset)
        d.max_push = float(bm.max_push)
        d.resolve = int(bool(bm.resolve))
        d.coherent = int(bm.coherent)
        self.ctx.call("xpbd_set_bone_mesh", C.byref(d))
        self._bm_nbv = d.nverts

    def set_bone_mesh_points(self, verts) -> None:
        v = f32(verts, (-1, 3))
        if v.shape[0] != getattr(self, "_bm_nbv", -1):
            raise ValueError(f"bone mesh has {getattr(self, '_bm_nbv', 0)} vertices, got {v.shape[0]}")
        self.ctx.call("xpbd_set_bone_mesh_points", ptr(v))

    def set_bone_mesh_resolve(self, on: bool) -> None:
        self.ctx.call("xpbd_set_bone_mesh_resolve", int(bool(on)))

    def bone_mesh_push(self) -> None:
        self.ctx.call("xpbd_bone_mesh_push")

    def bone_mesh_queries(self) -> int:
        """closest-point queries made so far (vertices not certified outside)"""
        n = np.zeros(1, np.int64)
        self.ctx.call("xpbd_bone_mesh_queries", ptr(n))
        return int(n[0])

    def bone_penetration(self, collide_eps: float = 1e-6, want_depth: bool = True):
        """(depth per vertex or None, count deeper than collide_eps, largest depth)"""
        depth = np.zeros(self.NV, np.float32) if want_depth else None
        cnt = C.c_int32(0)
        mx = C.c_float(0.0)
        self.ctx.call("xpbd_bone_penetration", C.c_float(collide_eps), ptr(depth), C.byref(cnt), C.byref(mx))
        return depth, int(cnt.value), float(mx.value)

    # fascia shell, ride mode (Fascia, src/muscle/fascia/fascia.py:45-572) -----------
    def set_fascia(self, fascia, link) -> None:
        """Link the finalized fascia engine (None unlinks). ``link``: dict with iters and
        mode 0 (ride: surf_tris, bind_max_dist) or 1 (squeeze: follow0, glue_fascia,
        glue_muscle, glue_restlen, glue_compliance, omega, anchor, anchor_offset)."""
        if fascia is None:
            self.ctx.call("xpbd_set_fascia", None, None)
            self._fascia = None
            return
        k = self._keep
        d = capi.XpbdFasciaDesc()
        d.iters = int(link["iters"])
        d.mode = int(link.get("mode", 0))
        if d.mode == 0:
            k["fascia_tris"] = i32(link["surf_tris"], (-1, 3))
            d.nsurf = k["fascia_tris"].shape[0]
            d.surf_tris = ptr(k["fascia_tris"])
            d.bind_max_dist = float(link["bind_max_dist"])
        else:
            arr = dict(follow0=f32(link["follow0"], (-1, 3)), glue_fascia=i32(link["glue_fascia"]),
                       glue_muscle=i32(link["glue_muscle"]), glue_restlen=f32(link["glue_restlen"]),
                       glue_compliance=f32(link["glue_compliance"]), anchor=i32(link["anchor"]),
                       anchor_offset=f32(link["anchor_offset"], (-1, 3)))
            k["fascia_squeeze"] = arr
            for name, a in arr.items():
                setattr(d, name, ptr(a))
            d.nglue = len(arr["glue_fascia"])
            d.nanchor = len(arr["anchor"])
            d.omega = float(link["omega"])
        self.ctx.call("xpbd_set_fascia", fascia.ctx.h, C.byref(d))
        self._fascia = fascia

    def set_fascia_follow(self, xyz) -> None:
        """squeeze: this frame's follow positions; advects the shell by their change"""
        a = f32(xyz, (-1, 3))
        self.ctx.call("xpbd_set_fascia_follow", ptr(a))

    def fascia_substep(self) -> None:
        self.ctx.call("xpbd_fascia_substep")

    def fascia_bind(self):
        """(face, weights (n, 3), gap, n bound) of the bind, per fascia vertex"""
        n = self._fascia.NV
        face = np.zeros(n, np.int32)
        w = np.zeros((n, 3), np.float32)
        gap = np.zeros(n, np.float32)
        nb = np.zeros(1, np.int32)
        self.ctx.call("xpbd_get_fascia_bind", ptr(face), ptr(w), ptr(gap), ptr(nb))
        return face, w, gap, int(nb[0])

    def _set_dcd(self, spec):
        from .collision import DcdParams
        p = spec.params or DcdParams()
        q = capi.XpbdDcdParams()
        for f, _ in capi.XpbdDcdParams._fields_:
            setattr(q, f, type(getattr(q, f))(getattr(p, f)))  # float or int field
        self._dcd_max_pairs = int(p.max_pairs) if p.max_pairs > 0 else (1 << 22)
        k = self._keep
        k["dcd_tris"] = i32(spec.tris, (-1, 3))
        k["dcd_group"] = None if spec.group is None else i32(spec.group)
        self.ctx.call("xpbd_set_dcd", k["dcd_tris"].shape[0], ptr(k["dcd_tris"]), ptr(k["dcd_group"]),
                      C.byref(q))

    def dcd_solve(self, rebuild: bool):
        self.ctx.call("xpbd_dcd_solve", int(bool(rebuild)))

    def dcd_pairs(self) -> np.ndarray:
        """The (i<j)-sorted pair list of the last broadphase (at most max_pairs
        of them; the stats give the number found)."""
        n = np.zeros(1, np.int64)
        self.ctx.call("xpbd_dcd_pairs", None, 0, ptr(n))
        m = min(int(n[0]), self._dcd_max_pairs)
        out = np.zeros((m, 2), np.int32)
        self.ctx.call("xpbd_dcd_pairs", ptr(out), m, ptr(n))
        return out

    def dcd_stats(self) -> dict:
        a = np.zeros(1, np.int64)
        b = np.zeros(1, np.int64)
        cell = np.zeros(1, np.float32)
        self.ctx.call("xpbd_dcd_stats", ptr(a), ptr(b), ptr(cell))
        lo = np.zeros(1, np.int64)
        tr = np.zeros(1, np.int64)
        self.ctx.call("xpbd_dcd_overflow", ptr(lo), ptr(tr))
        return {"n_pairs": int(a[0]), "n_contacts": int(b[0]), "cell": float(cell[0]),
                "left_out_tris": int(lo[0]), "truncated_tris": int(tr[0])}

    def set_params(self, p):
        q = capi.XpbdParams()
        q.dt = float(p.dt)
        for i in range(3):
            q.gravity[i] = float(p.gravity[i])
        q.veldamping = float(p.veldamping)
        q.quasi_static = int(bool(p.quasi_static))
        q.contraction_ratio = float(p.contraction_ratio)
        q.fiber_stiffness_scale = float(p.fiber_stiffness_scale)
        q.has_compress = int(bool(p.has_compress))
        q.exp_damp = float(p.exp_damp)
        q.const_damp = float(p.const_damp)
        self.ctx.call("xpbd_set_params", C.byref(q))
        self.dt = float(p.dt)

    def stats(self) -> dict:
        st = capi.XpbdStats()
        self.ctx.call("xpbd_get_stats", C.byref(st))
        return {f: getattr(st, f) for f, _ in capi.XpbdStats._fields_}

    # -- per-frame inputs ---------------------------------------------------
    def set_pin_targets(self, xyz):
        a = f32(xyz, (-1, 3))
        self._keep["pin_tgt"] = a
        self.ctx.call("xpbd_set_pin_targets", a.shape[0], ptr(a))

    def set_fiber_activation(self, act):
        """MultiGridSim.set_fiber_activation (mg.py:1098-1111): per-tet activation."""
        a = f32(act)
        if a.shape[0] != self.NT:
            raise ValueError(f"activation has {a.shape[0]} entries, expected NT={self.NT}")
        self._keep["act"] = a
        self.ctx.call("xpbd_set_activation", ptr(a))

    def set_activation_groups(self, tet_group):
        """one activation per group of tets (e.g. per batched instance), set once"""
        g = i32(tet_group)
        if g.shape[0] != self.NT:
            raise ValueError(f"tet_group has {g.shape[0]} entries, expected NT={self.NT}")
        self._keep["act_group"] = g
        self._n_act_groups = int(g.max()) + 1 if g.size else 0
        self.ctx.call("xpbd_set_activation_groups", self._n_act_groups, ptr(g))

    def set_group_activation(self, act):
        """per-group activation, expanded to the tets on the device"""
        a = f32(act)
        self._keep["gact"] = a
        self.ctx.call("xpbd_set_group_activation", a.shape[0], ptr(a))

    def set_bone_transforms(self, xforms):
        x = f32(xforms, (-1, 12))
        self._keep["xf"] = x
        self.ctx.call("xpbd_set_bone_transforms", x.shape[0], ptr(x))

    def set_attach_targets(self, xyz):
        a = f32(xyz, (-1, 3))
        self._keep["att"] = a
        self.ctx.call("xpbd_set_attach_targets", a.shape[0], ptr(a))

    # -- XpbdEngine primitives (engine.py:798-1134) -------------------------
    def integrate(self):
        self.ctx.call("xpbd_integrate")

    def clear(self):
        self.ctx.call("xpbd_clear")

    def solve_constraints(self, run_extras: bool = True):
        if not self.use_jacobi:
            raise XpbdError(capi.XPBD_EUNSUPPORTED, "Gauss-Seidel mode is not on the GPU path")
        self.ctx.call("xpbd_solve_constraints")
        if run_extras:
            for fn in self.constraint_solvers:
                fn()

    def apply_dP(self):
        self.ctx.call("xpbd_apply_dP")

    def apply_pintotarget(self):
        self.ctx.call("xpbd_apply_pins")

    def solve_extras(self):
        for fn in self.constraint_solvers:
            fn()

    def post_smooth(self, n_iters: int = 0):
        """post_smooth (engine.py:1072-1100): n Jacobi passes of clear/solve/apply/pin."""
        if n_iters > 0:
            self.ctx.call("xpbd_iterate", int(n_iters))

    def iterate(self, n: int):
        """n fused Jacobi iterations: clear, dispatch, apply_dP, pin (mg.py:1341-1360)."""
        self.ctx.call("xpbd_iterate", int(n))

    def collide(self):
        self.ctx.call("xpbd_collide")

    def update_velocities(self):
        self.ctx.call("xpbd_update_velocities")

    def native_substep(self, iters: int):
        self.ctx.call("xpbd_substep", int(iters))

    def native_frame(self, nsub: int, iters: int, use_graph: bool = True):
        self.ctx.call("xpbd_step_frame", int(nsub), int(iters), int(bool(use_graph)))

    def synchronize(self):
        self.ctx.call("xpbd_synchronize")

    # -- state ------------------------------------------------------------
    def get_state(self):
        n = self.NV
        pos = np.empty((n, 3), np.float32)
        pp = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        self.ctx.call("xpbd_get_state", ptr(pos), ptr(pp), ptr(vel))
        return pos, pp, vel

    def positions_np(self) -> np.ndarray:
        n = self.NV
        pos = np.empty((n, 3), np.float32)
        self.ctx.call("xpbd_get_state", ptr(pos), None, None)
        return pos

    def velocities_np(self) -> np.ndarray:
        n = self.NV
        vel = np.empty((n, 3), np.float32)
        self.ctx.call("xpbd_get_state", None, None, ptr(vel))
        return vel

    def set_state(self, pos=None, pprev=None, vel=None):
        a = None if pos is None else f32(pos, (-1, 3))
        b = None if pprev is None else f32(pprev, (-1, 3))
        v = None if vel is None else f32(vel, (-1, 3))
        self.ctx.call("xpbd_set_state", ptr(a), ptr(b), ptr(v))

    def get_rest(self):
        rv = np.empty(self.NT, np.float32)
        rm = np.empty((self.NT, 3, 3), np.float32)
        m = np.empty(self.NV, np.float32)
        self.ctx.call("xpbd_get_rest", ptr(rv), ptr(rm), ptr(m))
        return rv, rm, m

    # -- diagnostics --------------------------------------------------------
    def residuals(self) -> dict:
        """Per-type residual RMS (arap_residual_accum_kernel, mg/kernel.py:87-118,
        generalised per constraint type)."""
        out = np.zeros(8, np.float64)
        self.ctx.call("xpbd_residuals", ptr(out))
        r = {}
        for i, name in enumerate(("arap", "volume", "snh", "aniso")):
            cnt = out[2 * i + 1]
            if cnt > 0:
                r[name] = float(np.sqrt(out[2 * i] / cnt))
        return r

    def count_inverted(self) -> int:
        out = np.zeros(1, np.int32)
        self.ctx.call("xpbd_count_inverted", ptr(out))
        return int(out[0])

    def sdf_cells(self) -> np.ndarray:
        out = np.empty((self.NV, 2), np.int32)
        self.ctx.call("xpbd_get_sdf_cells", ptr(out))
        return out

    # -- two-level correction --------------------------------------------------
    def set_coarse_level(self, coarse: "XpbdEngine", level, kinematic: bool = True):
        """Link a coarse-cage engine (flat strategy) with the transfer tables of a
        multigrid.CoarseLevel."""
        k = self._keep
        arrs = dict(p_idx=i32(level.p_idx, (-1, 8)), p_w=f32(level.p_w, (-1, 8)), blk_off=i32(level.blk_off),
                    minv=f32(level.minv), eps=f32(level.eps), fit_off=i32(level.fit_off),
                    fit_fine=i32(level.fit_fine), fit_center=f32(level.fit_center, (-1, 3)),
                    fit_sinv=f32(level.fit_sinv, (-1, 16)))
        k["mg"] = arrs
        k["mg_coarse"] = coarse       # the coarse context must outlive this one's use of it
        d = capi.XpbdCoarseDesc(nblk=len(level.blk_off) - 1, iters=int(level.iters), kinematic=int(bool(kinematic)),
                                pin_rule=int(getattr(level, "pin_rule", 0)),
                                **{n: a.ctypes.data for n, a in arrs.items()})
        self.ctx.call("xpbd_set_coarse_level", coarse.ctx.h, C.byref(d))

    def set_coarse_pin_targets(self, xyz):
        """pin rule 1: the restricted bone goal of every coarse node (coarse order)"""
        a = f32(xyz, (-1, 3))
        self._keep["mg_pin_tgt"] = a
        self.ctx.call("xpbd_set_coarse_pin_targets", a.shape[0], ptr(a))

    def vcycle(self):
        self.ctx.call("xpbd_vcycle")

    # -- learned warm start ----------------------------------------------------
    def set_lambda_mode(self, mode: int):
        self.ctx.call("xpbd_set_lambda_mode", int(mode))

    def warm_start(self):
        self.ctx.call("xpbd_warm_start")

    # Device tensors exchanged with torch: torch's stream is drained before the
    # library reads one, and the library's stream before torch reads one.
    def _torch_sync(self):
        import torch
        torch.cuda.current_stream().synchronize()

    def features(self, out, normalized: bool = False):
        """per-element features into a (n_elems, 24) float32 CUDA tensor"""
        self._torch_sync()
        self.ctx.call("xpbd_features", C.c_void_p(out.data_ptr()), int(bool(normalized)))
        self.synchronize()

    def get_lambda(self, out):
        """per-element multipliers into a (n_elems, 4) float32 CUDA tensor"""
        self._torch_sync()
        self.ctx.call("xpbd_get_lambda", C.c_void_p(out.data_ptr()))
        self.synchronize()

    def set_lambda(self, t):
        self._torch_sync()
        self.ctx.call("xpbd_set_lambda", C.c_void_p(t.data_ptr()))
        self.synchronize()

    def set_reference_iters(self, n: int):
        self.ctx.call("xpbd_set_reference_iters", int(n))

    def set_warm_schedule(self, ref_iters: int, ref_every: int = 0, use_network: bool = True):
        """mode 4: a reference solve every ``ref_every``-th substep (0: first of a frame)."""
        self.ctx.call("xpbd_set_warm_schedule", int(ref_iters), int(ref_every), int(bool(use_network)))

    def snapshot(self, save: bool):
        self.ctx.call("xpbd_snapshot", int(bool(save)))

    def set_mlp(self, p: dict, feat_mean, feat_invstd, out_scale):
        k = self._keep
        k["mlp"] = [np.ascontiguousarray(p[n], np.float32) for n in ("W1", "b1", "W2", "b2", "W3", "b3")]
        k["mlp_norm"] = [np.ascontiguousarray(a, np.float32) for a in (feat_mean, feat_invstd, out_scale)]
        self.ctx.call("xpbd_set_mlp", *[ptr(a) for a in k["mlp"]], *[ptr(a) for a in k["mlp_norm"]])

    def kernel_timing(self, enable: bool):
        self.ctx.call("xpbd_kernel_timing", int(bool(enable)))

    def get_kernel_timing(self):
        ms = np.zeros(1, np.float64)
        n = np.zeros(1, np.int32)
        self.ctx.call("xpbd_get_kernel_timing", ptr(ms), ptr(n))
        return float(ms[0]), int(n[0])

    def time_iterations(self, n: int = 10, reps: int = 20) -> float:
        """n * reps projection iterations in one HIP graph, timed with HIP events: us per iteration."""
        us = np.zeros(1, np.float64)
        self.ctx.call("xpbd_time_iterations", int(n), int(reps), ptr(us))
        return float(us[0])

    def close(self):
        self.ctx.close()
