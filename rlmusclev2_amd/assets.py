"""The reference's asset-driven inputs of MultiGridSim (mg.py:95-576, 1113-1249).

``from_assets(cfg)`` builds a workload from the files a humanV2-style config names
instead of a ``synthetic`` block:

* ``rest_path`` (default ``<geo_dir>/restpos.geo``): the rest tet mesh, with the
  point attributes ``materialW`` (fibre direction), ``tendonmask`` and ``pinmask``
  when present (mg.py:299-347).  A USD rest (``.usd``/``.usda``/``.usdc``/``.usdz``)
  needs the ``pxr`` module, which this environment does not have: it raises.
* the per-frame targets, by ``target_mode`` (mg.py:424-458):

  - ``"points"``: the constrained vertices ``pts`` and their ``target_pos`` from
    ``<cons_dir>/cons_<start_frame>.geo``; ``rest_offset = rest[pts] - target_pos``;
    each frame's goal is ``cons_<frame>.geo``'s ``target_pos + rest_offset``
    (``ConsTargets``, load_frame_targets mg.py:1233-1249).
  - ``"mesh"``: every vertex is bound to the closest point of the bone surface of
    ``<bone_dir>/bone_<start_frame>.geo`` (``xpbd_mesh_bind``), its rest offset kept
    in the bound face's frame; each frame's goal rides that face through the
    ``bone_<frame>.geo`` sequence (``BoneSurfaceTargets``, _build_bone_slide_binding
    and bone_slide_target_kernel without re-projection, mg.py:1117-1176,
    mg/kernel.py:268-347).  ``use_sliding`` raises.

* the hard-pin subset by the ``muscleToBonePin`` region rule (``pin_active``,
  mg.py:486-536); the belly attach then covers the constrained vertices that are
  not hard-pinned (_build_fine_belly_attach, mg.py:700-734).
* ``bone_dir`` also gives the exact-mesh bone collider its per-frame skeleton
  (``geo.GeoBoneSource``).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import capi, geo
from .workloads import Workload

USD_SUFFIXES = (".usd", ".usda", ".usdc", ".usdz")


def mesh_bind(verts, tris, pts, max_dist: float):
    """(face, barycentric weights (n, 3), signed normal gap) of the closest point of
    each point on the triangle mesh among the faces nearer than ``max_dist``
    (face -1: none), from the library's host query (``xpbd_mesh_bind``)."""
    lib = capi.load_library()
    v = capi.f32(verts, (-1, 3))
    t = capi.i32(tris, (-1, 3))
    p = capi.f32(pts, (-1, 3))
    n = p.shape[0]
    face = np.empty(n, np.int32)
    w = np.empty((n, 3), np.float32)
    gap = np.empty(n, np.float32)
    rc = lib.xpbd_mesh_bind(v.shape[0], capi.ptr(v), t.shape[0], capi.ptr(t), n, capi.ptr(p),
                            C.c_float(max_dist), capi.ptr(face), capi.ptr(w), capi.ptr(gap))
    if rc != capi.XPBD_OK:
        raise capi.XpbdError(rc, "xpbd_mesh_bind: bad mesh, points or max_dist")
    return face, w, gap


def face_frames(corners: np.ndarray) -> np.ndarray:
    """Orthonormal frame per face, as (n, 3, 3) matrices whose columns are (t1, t2, n)
    (_face_basis, mg/kernel.py:258-285): with the corners p0, p1, p2 that Warp's
    mesh_eval_position gives at barycentric (0, 0), (1, 0), (0, 1) -- the third,
    first and second corner of the face under its u p_a + v p_b + (1 - u - v) p_c
    convention -- t1 = (p1 - p0)/|p1 - p0|, n = (p1 - p0) x (p2 - p0) normalised,
    t2 = n x t1.  FP32, as the reference's kernel."""
    c = np.asarray(corners, np.float32)
    p0, p1, p2 = c[:, 2], c[:, 0], c[:, 1]
    e1 = p1 - p0
    n = np.cross(e1, p2 - p0)
    n = n / np.linalg.norm(n, axis=1, keepdims=True).astype(np.float32)
    t1 = e1 / np.linalg.norm(e1, axis=1, keepdims=True).astype(np.float32)
    t2 = np.cross(n, t1)
    return np.stack([t1, t2, n], axis=2).astype(np.float32)


class ConsTargets:
    """target_mode "points": goal of constraint i = target_pos_f[i] + rest_offset[i]."""

    def __init__(self, cons_dir, start_frame: int, rest_pos: np.ndarray):
        self.cons_dir = str(cons_dir)
        c0 = geo.GeoFile(str(Path(cons_dir) / f"cons_{int(start_frame)}.geo"))
        self.pts = np.asarray(c0.pts, np.int32)
        self.rest_offset = (np.asarray(rest_pos, np.float32)[self.pts] - c0.target_pos).astype(np.float32)

    def goals(self, frame: int) -> np.ndarray:
        t = geo.read_cons_frame(self.cons_dir, int(frame))
        return (np.asarray(t, np.float32) + self.rest_offset).astype(np.float32)


class BoneSurfaceTargets:
    """target_mode "mesh" without sliding: every vertex rides its bound bone face."""

    def __init__(self, bone_source, start_frame: int, rest_pos: np.ndarray, max_dist: float):
        self.src = bone_source
        self.tris = np.asarray(bone_source.tris, np.int32).reshape(-1, 3)
        rest = np.asarray(rest_pos, np.float32)
        self.pts = np.arange(rest.shape[0], dtype=np.int32)
        bpos = np.asarray(bone_source.positions(int(start_frame)), np.float32)
        face, w, _ = mesh_bind(bpos, self.tris, rest, max_dist)
        # unbound vertices keep face 0 with zero weights and a zero offset, as the
        # reference's bind leaves them (bone_bind_kernel's else branch)
        self.ok = face >= 0
        self.face = np.where(self.ok, face, 0).astype(np.int32)
        self.w = np.where(self.ok[:, None], w, 0.0).astype(np.float32)
        corners = bpos[self.tris[self.face]]
        cp = np.einsum("nk,nkd->nd", self.w, corners).astype(np.float32)
        R = face_frames(corners)
        self.local = np.where(self.ok[:, None], np.einsum("nji,nj->ni", R, rest - cp), 0.0).astype(np.float32)

    def goals(self, frame: int) -> np.ndarray:
        bpos = np.asarray(self.src.positions(int(frame)), np.float32)
        corners = bpos[self.tris[self.face]]
        cp = np.einsum("nk,nkd->nd", self.w, corners).astype(np.float32)
        R = face_frames(corners)
        return (cp + np.einsum("nij,nj->ni", R, self.local)).astype(np.float32)


def _path(cfg: dict, key: str, default, base: Path) -> Path:
    p = Path(str(cfg.get(key, default)))
    return p if p.is_absolute() else base / p


def from_assets(cfg: dict, base=None) -> Workload:
    """Workload of an asset-driven config (paths relative to ``base``, default the
    working directory, as the reference resolves them)."""
    root = Path(base) if base is not None else Path.cwd()
    geo_dir = _path(cfg, "geo_dir", ".", root)
    rest_path = _path(cfg, "rest_path", geo_dir / "restpos.geo", root)
    if rest_path.suffix.lower() in USD_SUFFIXES:
        raise NotImplementedError(f"{rest_path}: a USD rest mesh (UsdTetRest, mesh_io.py:371-600) needs the pxr "
                                  "module, which is not installed; convert the asset to .geo")
    rest = geo.GeoFile(str(rest_path))
    pos0 = np.asarray(rest.positions, np.float32)
    tets = np.asarray(rest.tet_indices, np.int32).reshape(-1, 4)
    attrs = {k: np.asarray(rest.point_attr(k), np.float32) for k in ("tendonmask", "pinmask")
             if rest.has_point_attr(k)}
    fiber = np.asarray(rest.point_attr("materialW"), np.float32) if rest.has_point_attr("materialW") else None
    start = int(cfg.get("start_frame", 1))
    mode = str(cfg.get("target_mode", "points"))
    bone_src = None
    if cfg.get("bone_dir"):
        bone_src = geo.GeoBoneSource(_path(cfg, "bone_dir", "", root), start_frame=start,
                                     scale=float(cfg.get("asset_scale", 1.0)))
    if mode == "points":
        targets = ConsTargets(_path(cfg, "cons_dir", geo_dir, root), start, pos0)
    elif mode == "mesh":
        if bool(cfg.get("use_sliding", False)):
            raise NotImplementedError("use_sliding: re-projecting the bone binding each slide_interval frames "
                                      "is not on this path")
        if bone_src is None:
            raise ValueError("target_mode 'mesh' needs bone_dir")
        targets = BoneSurfaceTargets(bone_src, start, pos0, float(cfg.get("slide_range", 0.02)) + 1.0)
    else:
        raise NotImplementedError(f"target_mode={mode!r}: 'points' or 'mesh'")
    info = {"assets": dict(targets=targets, point_attrs=attrs, bone_source=bone_src, start_frame=start,
                           rest_path=str(rest_path))}
    if rest.has_prim_attr("muscle_id"):
        info["assets"]["muscle_id"] = (np.asarray(rest.prim_attr("muscle_id")), rest.prim_attr_strings("muscle_id"))
    return Workload(name="assets", pos0=pos0, tets=tets, v_fiber_dir=fiber,
                    v_tendonmask=attrs.get("tendonmask"), pin_vids=targets.pts, info=info)


def pin_active(cfg_pin: dict | None, w: Workload) -> np.ndarray:
    """Hard-pin subset of an asset workload's constrained vertices by the
    muscleToBonePin region rule (mg.py:486-536)."""
    pts = w.pin_vids
    attrs = w.info["assets"]["point_attrs"]
    if cfg_pin is None:
        return np.zeros(len(pts), np.int32)
    region = str(cfg_pin.get("region", "tendonmask")).lower()
    if region == "full":
        return np.ones(len(pts), np.int32)
    if region in ("tendonmask", "pinmask"):
        thr = cfg_pin.get("tendon_pin_threshold") if region == "tendonmask" else cfg_pin.get("pin_threshold", 0.5)
        if thr is None:
            raise ValueError(f"muscleToBonePin region='{region}' requires "
                             f"{'tendon_pin_threshold' if region == 'tendonmask' else 'pin_threshold'}")
        if region not in attrs:
            if region == "pinmask":        # no authored pinmask: every vertex rides the belly attach
                return np.zeros(len(pts), np.int32)
            raise ValueError(f"muscleToBonePin region='{region}' needs a '{region}' point attr on the rest asset")
        return (attrs[region][pts] > float(thr)).astype(np.int32)
    raise ValueError(f"muscleToBonePin 'region' must be 'full', 'tendonmask' or 'pinmask', got {region!r}")
