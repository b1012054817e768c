"""Constraint table construction (vectorised NumPy).

Produces the struct-of-arrays constraint table that the reference uploads in
``XpbdEngine._upload_cons`` (src/muscle/engine.py:1179-1211), with the same
per-type rules as ``ConstraintBuilderMixin`` (src/muscle/constraint/builder.py):

* rest data in float64: Dm columns ``p_k - p3``, signed volume ``det/6``,
  a tet is valid iff ``|det| > 1e-30`` (builder.py:28-43);
* volume / SNH / ARAP-SVD / aniso_arap: one constraint per valid tet with
  ``restlength`` = signed volume (builder.py:155-177, 310-345, 583-608);
* ATTACH: ``pts=[v,-1,-1,-1]``, target in ``restvector`` (mg.py:700-723);
* the table is stable-sorted by type and split into ``cons_ranges``
  (musclesim.py:369-389).

The reference builds a Python dict per constraint; at 535k tets that is slow,
so the same fields are produced here as whole arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .defs import ConstraintType


@dataclass
class ConsTable:
    type: np.ndarray
    pts: np.ndarray
    stiffness: np.ndarray
    compressionstiffness: np.ndarray
    dampingratio: np.ndarray
    tetid: np.ndarray
    L: np.ndarray
    restlength: np.ndarray
    restvector: np.ndarray
    restdir: np.ndarray
    cons_ranges: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    @staticmethod
    def empty() -> "ConsTable":
        return ConsTable(
            type=np.zeros(0, np.int32), pts=np.zeros((0, 4), np.int32),
            stiffness=np.zeros(0, np.float32), compressionstiffness=np.zeros(0, np.float32),
            dampingratio=np.zeros(0, np.float32), tetid=np.zeros(0, np.int32),
            L=np.zeros((0, 3), np.float32), restlength=np.zeros(0, np.float32),
            restvector=np.zeros((0, 4), np.float32), restdir=np.zeros((0, 3), np.float32))

    @staticmethod
    def concat(parts) -> "ConsTable":
        parts = [p for p in parts if p is not None and p.n > 0]
        if not parts:
            return ConsTable.empty()
        f = lambda k: np.concatenate([getattr(p, k) for p in parts], axis=0)
        return ConsTable(type=f("type"), pts=f("pts"), stiffness=f("stiffness"),
                         compressionstiffness=f("compressionstiffness"),
                         dampingratio=f("dampingratio"), tetid=f("tetid"), L=f("L"),
                         restlength=f("restlength"), restvector=f("restvector"),
                         restdir=f("restdir"))

    def sorted_by_type(self) -> "ConsTable":
        """Stable sort by type tag + cons_ranges {type: (offset, count)} in
        ascending tag order (musclesim.py:369-389)."""
        order = np.argsort(self.type, kind="stable")
        t = ConsTable(type=self.type[order], pts=self.pts[order], stiffness=self.stiffness[order],
                      compressionstiffness=self.compressionstiffness[order],
                      dampingratio=self.dampingratio[order], tetid=self.tetid[order],
                      L=self.L[order], restlength=self.restlength[order],
                      restvector=self.restvector[order], restdir=self.restdir[order])
        ranges = {}
        if t.n:
            vals, starts, counts = np.unique(t.type, return_index=True, return_counts=True)
            for v, s, c in zip(vals, starts, counts):
                ranges[ConstraintType(int(v))] = (int(s), int(c))
        t.cons_ranges = ranges
        return t

    def contiguous(self) -> "ConsTable":
        for k in ("type", "pts", "tetid"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.int32))
        for k in ("stiffness", "compressionstiffness", "dampingratio", "L", "restlength",
                  "restvector", "restdir"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float32))
        return self


def tet_rest_data(pos0: np.ndarray, tets: np.ndarray):
    """float64 rest matrices / signed volumes / validity (builder.py:28-43)."""
    tp = pos0.astype(np.float64)[tets]
    cols = tp[:, :3, :] - tp[:, 3:4, :]
    M = np.transpose(cols, (0, 2, 1))
    dets = np.linalg.det(M)
    volumes = dets / 6.0
    valid = np.abs(dets) > 1e-30
    return volumes, valid


def _tet_table(ctype, tets, valid, volumes, stiffness, dampingratio, restdir0=0.0,
               restvector=(0.0, 0.0, 0.0, 1.0)) -> ConsTable:
    idx = np.nonzero(valid)[0].astype(np.int32)
    n = idx.shape[0]
    restdir = np.zeros((n, 3), np.float32)
    restdir[:, 0] = restdir0
    return ConsTable(
        type=np.full(n, int(ctype), np.int32),
        pts=tets[idx].astype(np.int32),
        stiffness=np.full(n, stiffness, np.float32),
        compressionstiffness=np.full(n, -1.0, np.float32),
        dampingratio=np.full(n, dampingratio, np.float32),
        tetid=idx,
        L=np.zeros((n, 3), np.float32),
        restlength=volumes[idx].astype(np.float32),
        restvector=np.tile(np.asarray(restvector, np.float32), (n, 1)),
        restdir=restdir)


def create_tet_volume(tets, valid, volumes, params) -> ConsTable:
    """create_tet_volume_constraint (builder.py:155-177)."""
    return _tet_table(ConstraintType.TETVOLUME, tets, valid, volumes,
                      params.get("stiffness", 1e10), params.get("dampingratio", 0.0),
                      restvector=(0.0, 0.0, 0.0, 0.0))


def create_tet_arap(tets, valid, volumes, params, ctype=ConstraintType.TETARAPSVD) -> ConsTable:
    """create_tet_arap_constraints (builder.py:583-608); also aniso_arap (:630-635)."""
    return _tet_table(ctype, tets, valid, volumes, params.get("stiffness", 1e6),
                      params.get("dampingratio", 0.0))


def create_tet_snh(tets, valid, volumes, params) -> ConsTable:
    """create_tet_snh_constraint (builder.py:310-345): mu in stiffness, lam in restdir[0]."""
    return _tet_table(ConstraintType.TETSNH, tets, valid, volumes, params.get("mu", 1000.0),
                      params.get("dampingratio", 0.0), restdir0=params.get("lam", 10000.0))


def create_attach(vertices, targets, stiffness, dampingratio, restlength=0.0, target_ids=None,
                  tetid=None) -> ConsTable:
    """ATTACH per vertex: pts = [v, -1, target id or -1, -1], target in restvector
    (MultiGridSim._build_fine_belly_attach, mg.py:700-723; create_attach_constraints,
    constraint/builder.py:347-398)."""
    v = np.asarray(vertices, np.int32)
    n = v.shape[0]
    pts = np.full((n, 4), -1, np.int32)
    pts[:, 0] = v
    if target_ids is not None:
        pts[:, 2] = np.asarray(target_ids, np.int32)
    rv = np.zeros((n, 4), np.float32)
    rv[:, 3] = 1.0
    if targets is not None:
        rv[:, :3] = targets
    return ConsTable(
        type=np.full(n, int(ConstraintType.ATTACH), np.int32), pts=pts,
        stiffness=np.full(n, stiffness, np.float32),
        compressionstiffness=np.full(n, -1.0, np.float32),
        dampingratio=np.full(n, dampingratio, np.float32),
        tetid=np.full(n, -1, np.int32) if tetid is None else np.asarray(tetid, np.int32),
        L=np.zeros((n, 3), np.float32),
        restlength=np.broadcast_to(np.asarray(restlength, np.float32), (n,)).copy(),
        restvector=rv, restdir=np.zeros((n, 3), np.float32))


def create_pintotarget(vertices, targets, rest_offset) -> ConsTable:
    """PINTOTARGET (builder.py:400-520): pos[src] = restvector + restdir."""
    v = np.asarray(vertices, np.int32)
    n = v.shape[0]
    pts = np.full((n, 4), -1, np.int32)
    pts[:, 0] = v
    rv = np.zeros((n, 4), np.float32)
    rv[:, :3] = targets
    rv[:, 3] = 1.0
    return ConsTable(
        type=np.full(n, int(ConstraintType.PINTOTARGET), np.int32), pts=pts,
        stiffness=np.zeros(n, np.float32), compressionstiffness=np.full(n, -1.0, np.float32),
        dampingratio=np.zeros(n, np.float32), tetid=np.full(n, -1, np.int32),
        L=np.zeros((n, 3), np.float32), restlength=np.zeros(n, np.float32),
        restvector=rv, restdir=np.asarray(rest_offset, np.float32).reshape(n, 3))


def build_from_configs(pos0, tets, constraint_configs) -> ConsTable:
    """_collect_raw_constraints (builder.py:610-656) for the tet types on the HIP
    path, then the stable type sort."""
    volumes, valid = tet_rest_data(pos0, tets)
    parts = []
    for params in constraint_configs:
        ctype = str(params["type"]).lower()
        if ctype == "volume":
            parts.append(create_tet_volume(tets, valid, volumes, params))
        elif ctype == "tetarapsvd":
            parts.append(create_tet_arap(tets, valid, volumes, params))
        elif ctype == "aniso_arap":
            parts.append(create_tet_arap(tets, valid, volumes, params, ConstraintType.TETARAPANISO))
        elif ctype == "snh":
            parts.append(create_tet_snh(tets, valid, volumes, params))
        elif ctype in ("attach", "pintotarget"):
            raise NotImplementedError(
                f"constraint type {ctype!r} needs per-vertex targets: MultiGridSim builds ATTACH from its "
                "'belly_attach' entry, MuscleSim.from_procedural from bone_targets and a vertex mask "
                "(create_attach / create_pintotarget take explicit targets)")
        else:
            raise NotImplementedError(
                f"constraint type {ctype!r} is not on the HIP substep path "
                "(tet types here: tetarapsvd, volume, snh, aniso_arap)")
    return ConsTable.concat(parts).sorted_by_type().contiguous()


def vertex_fiber_from_tets(n_verts, tets, fiber_dirs_per_tet):
    """Per-tet fibre dirs averaged to vertices, normalised (musclesim.py:220-229)."""
    vf = np.zeros((n_verts, 3), np.float32)
    vc = np.zeros(n_verts, np.float32)
    f = np.asarray(fiber_dirs_per_tet, np.float32)
    for k in range(4):
        np.add.at(vf, tets[:, k], f)
        np.add.at(vc, tets[:, k], 1.0)
    vc = np.maximum(vc, 1.0)
    vf /= vc[:, None]
    vf /= np.maximum(np.linalg.norm(vf, axis=1, keepdims=True), 1e-8)
    return vf.astype(np.float32)


def tet_tendonmask(tets, v_tendonmask):
    """compute_cell_tendon_mask_kernel (src/muscle/engine.py:722-732): mean of 4."""
    vt = np.asarray(v_tendonmask, np.float32)
    s = vt[tets[:, 0]] + vt[tets[:, 1]]
    s = s + vt[tets[:, 2]]
    s = s + vt[tets[:, 3]]
    return (s / np.float32(4.0)).astype(np.float32)
