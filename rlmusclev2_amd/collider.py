"""The bone triangle-mesh collider and the penetration meter, on the HIP library.

``AnimatedBoneMeshCollider`` and ``CollisionMeter`` keep the reference's names,
arguments and methods (src/example/humanV2.py:140-199, 319-418).  The collider
installs the skeleton mesh in the sim's native engine (``xpbd_set_bone_mesh``);
placed in ``sim.collision_solvers`` it pushes inside every native substep as the
collision stage's last word, without breaking the frame's hipGraph.  The query
and the push are the library's (closest point within ``max_dist``, the sign of
the closest feature's angle-weighted pseudonormal, push to ``contact_offset``).
"""
from __future__ import annotations

import numpy as np

from . import workloads
from .scene import BoneMeshCollider


class CapsuleBoneSource:
    """Bone source of a synthetic workload, the counterpart of the reference's
    ``make_bone_source`` (a ``.geo`` directory or an animated USD): constant
    triangles, per-frame vertex positions.  Triangles come in the winding of
    Houdini geometry (clockwise seen from outside), as the reference's SNU bone
    files do, which is why the collider flips them by default."""

    def __init__(self, w, n_around: int = 96, n_cap: int = 16, n_len: int = 96):
        self.w = w
        self.res = dict(n_around=n_around, n_cap=n_cap, n_len=n_len)
        verts, tris, toff, _ = workloads.bone_meshes(w, **self.res)
        self.n_verts = int(verts.shape[0])
        self.tris = np.ascontiguousarray(tris[:, [0, 2, 1]]).reshape(-1)
        self.tri_off = toff

    def positions(self, frame) -> np.ndarray:
        return workloads.bone_mesh_points(self.w, frame, **self.res)


class AnimatedBoneMeshCollider:
    """Per-frame skeleton mesh as a moving rigid collider (humanV2.py:319-418)."""

    native_stage = True     # runs inside the native substep when listed in sim.collision_solvers

    def __init__(self, sim, bone_source, max_dist: float = 0.2, contact_offset: float = 0.0,
                 max_push: float = -1.0, flip_winding: bool = True):
        self.sim = sim
        self.source = bone_source
        self.start_frame = int(getattr(sim, "start_frame", 1))
        self.max_dist = float(max_dist)
        self.contact_offset = float(contact_offset)
        self.max_push = float(max_push)
        tri = np.asarray(bone_source.tris, np.int32).reshape(-1, 3)
        if flip_winding:
            tri = tri[:, [0, 2, 1]]
        self.tris = np.ascontiguousarray(tri)
        self.n_bone_verts = int(bone_source.n_verts)
        self.n_bone_tris = int(tri.shape[0])
        nv = int(sim.NV)
        # fascia (is_fascia of the workload, the reference's verts >= nvm) is neither
        # pushed nor metered (humanV2.py:354-375)
        coll = np.ones(nv, np.float32)
        w = getattr(sim, "workload", None)
        if w is not None and w.is_fascia is not None:
            coll[np.asarray(w.is_fascia, bool)] = 0.0
        verts0 = bone_source.positions(self.start_frame)
        self._cur_verts = verts0
        self._bone_stack = None
        sim.engine.set_bone_mesh(BoneMeshCollider(
            verts=verts0, tris=self.tris, tri_off=getattr(bone_source, "tri_off", None), max_dist=self.max_dist,
            contact_offset=self.contact_offset, max_push=self.max_push, coll_w=coll, meter_w=coll.copy(),
            resolve=False))

    def preload(self, frames) -> None:
        self._bone_stack = {f: self.source.positions(f) for f in frames}

    def set_frame(self, frame) -> None:
        """This frame's bone pose into the mesh, then the refit."""
        if self._bone_stack is not None and frame in self._bone_stack:
            verts = self._bone_stack[frame]
        else:
            verts = self.source.positions(frame)
        if verts.shape[0] != self.n_bone_verts:
            raise ValueError(f"bone frame {frame} has {verts.shape[0]} verts, "
                             f"expected {self.n_bone_verts} (topology must be constant)")
        self.sim.engine.set_bone_mesh_points(verts)
        self._cur_verts = verts

    def __call__(self) -> None:
        """One push of the current positions (outside the native substep)."""
        self.sim.engine.bone_mesh_push()

    def inside_depth(self) -> np.ndarray:
        """Per-vertex penetration depth (0 outside, > 0 inside, metres)."""
        return self.sim.engine.bone_penetration(0.0, want_depth=True)[0]

    @property
    def tris_flat(self) -> np.ndarray:
        return self.tris.reshape(-1).astype(np.int32)


class CollisionMeter:
    """Per-frame penetration meter (humanV2.py:140-199): depth and 0/1 flags at
    ``collide_eps``; the bone meter only (no muscle-muscle meter on this path)."""

    def __init__(self, collider, mm_meter=None, collide_eps: float = 1e-6, NV: int | None = None):
        self.collider = collider
        self.mm_meter = mm_meter
        self.collide_eps = float(collide_eps)
        self.NV = NV
        self.bone_depth = self.bone_flag = None
        self.self_depth = self.self_flag = None

    def measure(self) -> "CollisionMeter":
        self.bone_depth = self.collider.inside_depth()
        self.bone_flag = (self.bone_depth > self.collide_eps).astype(np.int32)
        self.self_depth = self.self_flag = None
        if self.mm_meter is not None:
            dmm = self.mm_meter()
            self.self_depth = np.zeros(self.NV, dtype=np.float32)
            self.self_depth[:dmm.shape[0]] = dmm
            self.self_flag = (self.self_depth > self.collide_eps).astype(np.int32)
        return self

    def summary(self) -> str:
        mm = (f"mm={int(self.self_flag.sum())}({self.self_depth.max() * 1000:.1f}mm)"
              if self.self_flag is not None else "")
        return f"coll={int(self.bone_flag.sum())}({self.bone_depth.max() * 1000:.1f}mm) {mm}"
