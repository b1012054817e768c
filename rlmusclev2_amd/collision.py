"""Collision set-up on the host: surface triangles and DCD parameters.

``surface_tris`` follows vectorized_surface_tris (src/muscle/util/mesh_io.py:58-82):
a face is on the surface iff exactly one tet has it; the oriented face of
that tet is kept.  ``DcdParams`` mirrors UdcfParams
(src/muscle/collision/udcf/solver.py:37-105) and the humanV2 driver's
self-collision settings (src/example/humanV2.py:820-866).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FACE_SLOTS = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]], np.int32)


def surface_tris(tets: np.ndarray) -> np.ndarray:
    tets = np.asarray(tets)
    tris = tets[:, FACE_SLOTS].reshape(-1, 3)
    keys = np.sort(tris, axis=1)
    _, inv, counts = np.unique(keys, axis=0, return_inverse=True, return_counts=True)
    return tris[counts[inv.reshape(-1)] == 1].astype(np.int32)


@dataclass
class DcdParams:
    thickness: float = 1e-3
    multiplier: float = 0.3
    friction_static: float = 0.5
    friction_dynamic: float = 0.4
    offset: float = 1.0e-6
    normal_cull: float = 1.0
    cell_size: float = 0.0        # 0 = from the median triangle radius at each rebuild
    allow_self: bool = True
    iters: int = 2
    rebuild_substeps: bool = False
    max_pairs: int = 1 << 22


@dataclass
class DcdSpec:
    tris: np.ndarray
    group: np.ndarray | None = None
    params: DcdParams = None
