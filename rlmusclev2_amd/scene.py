"""Problem description shared by the HIP engine and its checker.

A ``Scene`` carries exactly the fields the reference sim objects hand to the
substep kernels: rest mesh, mass/density, the constraint table, fibre data,
engine scalars, kinematic pins and the bone SDF colliders
(src/muscle/musclesim.py:331-352, src/muscle/mg/mg.py:95-576).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .builder import ConsTable


@dataclass
class SimParams:
    dt: float                                   # substep dt
    gravity: tuple = (0.0, 0.0, 0.0)
    veldamping: float = 0.0
    quasi_static: bool = False
    contraction_ratio: float = 0.4
    fiber_stiffness_scale: float = 1.0
    has_compress: bool = False
    exp_damp: float = 1.0
    const_damp: float = 0.0


@dataclass
class SdfColliders:
    """Bone SDF grids (bone-local), rigid transforms and collision weights."""
    grids: np.ndarray          # (nbones, res, res, res) f32, index [b, z, y, x]
    origins: np.ndarray        # (nbones, 3) bone-local grid origin
    spacing: np.ndarray        # (nbones,) grid spacing
    contact_offset: float = 1e-5
    max_push: float = -1.0
    coll_w: np.ndarray | None = None          # (nv,) 0 = never pushed
    vertex_group: np.ndarray | None = None    # (nv,) bone group per vertex
    group_bone_off: np.ndarray | None = None  # (ngroups+1,)

    @property
    def nbones(self) -> int:
        return int(self.grids.shape[0])

    @property
    def res(self) -> int:
        return int(self.grids.shape[1])


@dataclass
class BoneMeshCollider:
    """Triangle mesh over the bones (AnimatedBoneMeshCollider, src/example/humanV2.py:319-384):
    one or more closed sub-meshes, outward winding; per-frame vertex positions, fixed topology."""
    verts: np.ndarray                         # (nbv, 3) positions of the first frame
    tris: np.ndarray                          # (ntris, 3) vertex ids, (b - a) x (c - a) outward
    tri_off: np.ndarray | None = None         # (nbones+1,) sub-mesh triangle ranges
    max_dist: float = 0.2
    contact_offset: float = 1e-5
    max_push: float = -1.0
    coll_w: np.ndarray | None = None          # (nv,) 0 = never pushed
    meter_w: np.ndarray | None = None         # (nv,) 0 = not metered
    vertex_group: np.ndarray | None = None    # (nv,) group per vertex
    group_bone_off: np.ndarray | None = None  # (ngroups+1,) sub-mesh ranges per group
    resolve: bool = True                      # False: penetration meter only
    coherent: int = 2                         # skip vertices certified outside: 0 off, 1 on, 2 when closed

    @property
    def nbones(self) -> int:
        return 1 if self.tri_off is None else int(len(self.tri_off) - 1)


@dataclass
class Scene:
    pos0: np.ndarray
    tets: np.ndarray
    density: float
    cons: ConsTable
    params: SimParams
    v_fiber_dir: np.ndarray | None = None
    tendonmask: np.ndarray | None = None       # per tet
    stopped: np.ndarray | None = None
    mass: np.ndarray | None = None             # overrides density-lumped mass
    pin_vids: np.ndarray | None = None
    pin_active: np.ndarray | None = None
    sdf: SdfColliders | None = None
    bone_mesh: BoneMeshCollider | None = None
    dcd: object | None = None                  # collision.DcdSpec
    attach_vids: np.ndarray | None = None      # vertex of each ATTACH constraint, table order
    rest_scale: float = 1.0                    # shape constraints see the rest shape scaled by this
    meta: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return int(self.pos0.shape[0])

    @property
    def nt(self) -> int:
        return int(self.tets.shape[0])
