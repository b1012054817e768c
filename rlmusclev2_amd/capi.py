"""ctypes binding of the HIP substep library (include/xpbd.h).

The library is built in-tree (``rlmusclev2_amd/_lib/libxpbd_hip.so``) by
``rlmusclev2_amd.build``.  There is no fallback: if the library is missing or
no GPU is visible the calls raise ``XpbdError``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libxpbd_hip.so"

XPBD_OK = 0
XPBD_EINVAL = 1
XPBD_EUNSUPPORTED = 2
XPBD_EOOM = 3
XPBD_EHIP = 4
XPBD_ELAYOUT = 5
XPBD_ENOTREADY = 6
_CODES = {1: "EINVAL", 2: "EUNSUPPORTED", 3: "EOOM", 4: "EHIP", 5: "ELAYOUT", 6: "ENOTREADY"}


class XpbdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"xpbd error {_CODES.get(code, code)}: {msg}")
        self.code = code


class XpbdParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("gravity", C.c_float * 3), ("veldamping", C.c_float),
                ("quasi_static", C.c_int32), ("contraction_ratio", C.c_float),
                ("fiber_stiffness_scale", C.c_float), ("has_compress", C.c_int32),
                ("exp_damp", C.c_float), ("const_damp", C.c_float)]


class XpbdTuning(C.Structure):
    _fields_ = [("patch_verts", C.c_int32), ("strategy", C.c_int32), ("block_tets", C.c_int32),
                ("block_threads", C.c_int32), ("resident_blocks", C.c_int32), ("reserved", C.c_int32 * 3)]


class XpbdCoarseDesc(C.Structure):
    _fields_ = [("p_idx", C.c_void_p), ("p_w", C.c_void_p), ("nblk", C.c_int32), ("blk_off", C.c_void_p),
                ("minv", C.c_void_p), ("eps", C.c_void_p), ("fit_off", C.c_void_p), ("fit_fine", C.c_void_p),
                ("fit_center", C.c_void_p), ("fit_sinv", C.c_void_p), ("iters", C.c_int32),
                ("kinematic", C.c_int32), ("pin_rule", C.c_int32)]


class XpbdDcdParams(C.Structure):
    _fields_ = [("thickness", C.c_float), ("multiplier", C.c_float), ("friction_static", C.c_float),
                ("friction_dynamic", C.c_float), ("offset", C.c_float), ("normal_cull", C.c_float),
                ("cell_size", C.c_float), ("allow_self", C.c_int32), ("iters", C.c_int32),
                ("rebuild_substeps", C.c_int32), ("max_pairs", C.c_int64)]


class XpbdBoneMeshDesc(C.Structure):
    _fields_ = [("nverts", C.c_int32), ("verts", C.c_void_p), ("ntris", C.c_int32), ("tris", C.c_void_p),
                ("nbones", C.c_int32), ("tri_off", C.c_void_p), ("ngroups", C.c_int32), ("goff", C.c_void_p),
                ("vgroup", C.c_void_p), ("coll_w", C.c_void_p), ("meter_w", C.c_void_p),
                ("max_dist", C.c_float), ("contact_offset", C.c_float), ("max_push", C.c_float),
                ("resolve", C.c_int32), ("coherent", C.c_int32)]


class XpbdFasciaDesc(C.Structure):
    _fields_ = [("nsurf", C.c_int32), ("surf_tris", C.c_void_p), ("bind_max_dist", C.c_float),
                ("iters", C.c_int32), ("mode", C.c_int32), ("follow0", C.c_void_p), ("nglue", C.c_int32),
                ("glue_fascia", C.c_void_p), ("glue_muscle", C.c_void_p), ("glue_restlen", C.c_void_p),
                ("glue_compliance", C.c_void_p), ("omega", C.c_float), ("nanchor", C.c_int32),
                ("anchor", C.c_void_p), ("anchor_offset", C.c_void_p)]


class XpbdMlpDev(C.Structure):
    _fields_ = [("W1", C.c_void_p), ("b1", C.c_void_p), ("W2", C.c_void_p), ("b2", C.c_void_p),
                ("W3", C.c_void_p), ("b3", C.c_void_p)]


class XpbdStats(C.Structure):
    _fields_ = [("nv", C.c_int32), ("nt", C.c_int32), ("n_elems", C.c_int32),
                ("n_patches", C.c_int32), ("n_patch_tets", C.c_int32), ("max_table", C.c_int32),
                ("max_patch_tets", C.c_int32), ("feature_mask", C.c_int32),
                ("bytes_per_iter_alg", C.c_double), ("bytes_per_iter_impl", C.c_double),
                ("resident_blocks", C.c_int32), ("resident_parts", C.c_int32), ("max_part_verts", C.c_int32),
                ("max_part_neighbours", C.c_int32), ("cross_vertex_entries", C.c_int32),
                ("resident_lds", C.c_int32)]


_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
# name -> argtypes (restype int32 unless listed in _RESTYPE)
SIGNATURES = {
    "xpbd_create": [_I, C.POINTER(_P)],
    "xpbd_destroy": [_P],
    "xpbd_last_error": [_P],
    "xpbd_set_stream": [_P, _P],
    "xpbd_set_mesh": [_P, _I, _P, _I, _P, _F],
    "xpbd_get_rest": [_P, _P, _P, _P],
    "xpbd_set_mass": [_P, _P],
    "xpbd_set_stopped": [_P, _P],
    "xpbd_set_force": [_P, _P],
    "xpbd_set_fiber": [_P, _P, _P],
    "xpbd_set_activation": [_P, _P],
    "xpbd_set_activation_groups": [_P, _I, _P],
    "xpbd_set_group_activation": [_P, _I, _P],
    "xpbd_upload_constraints": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "xpbd_set_attach_targets": [_P, _I, _P],
    "xpbd_set_params": [_P, C.POINTER(XpbdParams)],
    "xpbd_set_pins": [_P, _I, _P, _P],
    "xpbd_set_pin_targets": [_P, _I, _P],
    "xpbd_set_sdf": [_P, _I, _I, _P, _P, _P, _F, _F, _P],
    "xpbd_set_sdf_groups": [_P, _I, _P, _P],
    "xpbd_set_bone_transforms": [_P, _I, _P],
    "xpbd_finalize": [_P, C.POINTER(XpbdTuning)],
    "xpbd_get_stats": [_P, C.POINTER(XpbdStats)],
    "xpbd_mesh_bind": [_I, _P, _I, _P, _I, _P, C.c_float, _P, _P, _P],
    "xpbd_set_state": [_P, _P, _P, _P],
    "xpbd_get_state": [_P, _P, _P, _P],
    "xpbd_integrate": [_P],
    "xpbd_clear": [_P],
    "xpbd_solve_constraints": [_P],
    "xpbd_apply_dP": [_P],
    "xpbd_apply_pins": [_P],
    "xpbd_iterate": [_P, _I],
    "xpbd_collide": [_P],
    "xpbd_update_velocities": [_P],
    "xpbd_substep": [_P, _I],
    "xpbd_step_frame": [_P, _I, _I, _I],
    "xpbd_synchronize": [_P],
    "xpbd_residuals": [_P, _P],
    "xpbd_count_inverted": [_P, _P],
    "xpbd_get_sdf_cells": [_P, _P],
    "xpbd_set_dcd": [_P, _I, _P, _P, C.POINTER(XpbdDcdParams)],
    "xpbd_set_bone_mesh": [_P, C.POINTER(XpbdBoneMeshDesc)],
    "xpbd_set_bone_mesh_points": [_P, _P],
    "xpbd_set_bone_mesh_resolve": [_P, _I],
    "xpbd_bone_mesh_push": [_P],
    "xpbd_bone_penetration": [_P, _F, _P, _P, _P],
    "xpbd_bone_mesh_queries": [_P, _P],
    "xpbd_set_fascia": [_P, _P, C.POINTER(XpbdFasciaDesc)],
    "xpbd_get_fascia_bind": [_P, _P, _P, _P, _P],
    "xpbd_fascia_substep": [_P],
    "xpbd_set_fascia_follow": [_P, _P],
    "xpbd_scale_rest_shape": [_P, _F],
    "xpbd_dcd_solve": [_P, _I],
    "xpbd_dcd_pairs": [_P, _P, C.c_int64, _P],
    "xpbd_dcd_stats": [_P, _P, _P, _P],
    "xpbd_dcd_overflow": [_P, _P, _P],
    "xpbd_mlp_forward": [C.POINTER(XpbdMlpDev), _P, C.c_int64, _P, _P, _P, _P],
    "xpbd_mlp_grad_size": [],
    "xpbd_mlp_backward": [C.POINTER(XpbdMlpDev), _P, _P, _P, _P, C.c_int64, _P, _P, C.c_int64, _P],
    "xpbd_mlp_backward_noalloc": [C.POINTER(XpbdMlpDev), _P, _P, _P, _P, C.c_int64, _P, _P],
    "xpbd_set_coarse_level": [_P, _P, C.POINTER(XpbdCoarseDesc)],
    "xpbd_set_coarse_pin_targets": [_P, _I, _P],
    "xpbd_vcycle": [_P],
    "xpbd_set_lambda_mode": [_P, _I],
    "xpbd_set_mlp": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "xpbd_warm_start": [_P],
    "xpbd_features": [_P, _P, _I],
    "xpbd_get_lambda": [_P, _P],
    "xpbd_set_lambda": [_P, _P],
    "xpbd_set_reference_iters": [_P, _I],
    "xpbd_set_warm_schedule": [_P, _I, _I, _I],
    "xpbd_snapshot": [_P, _I],
    "xpbd_kernel_timing": [_P, _I],
    "xpbd_get_kernel_timing": [_P, _P, _P],
    "xpbd_time_iterations": [_P, _I, _I, _P],
}
_RESTYPE = {"xpbd_destroy": None, "xpbd_last_error": C.c_char_p}

_lib = None


def load_library(path: os.PathLike | None = None):
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise XpbdError(XPBD_EHIP, f"HIP library not built: {p} (run rlmusclev2_amd.build.build())")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int32)
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def f32(a, shape=None) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float32)
    return x.reshape(shape) if shape is not None else x


def i32(a, shape=None) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.int32)
    return x.reshape(shape) if shape is not None else x


class Context:
    """Owns one ``xpbd_ctx``; every method checks the status code."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.xpbd_create(int(device), C.byref(h))
        if rc != XPBD_OK:
            raise XpbdError(rc, f"xpbd_create(device={device}) failed (is a GPU visible?)")
        self.h = h
        self._keep = []

    def check(self, rc: int, what: str):
        if rc != XPBD_OK:
            msg = self.lib.xpbd_last_error(self.h)
            raise XpbdError(rc, f"{what}: {msg.decode() if msg else ''}")

    def call(self, name: str, *args):
        self.check(getattr(self.lib, name)(self.h, *args), name)

    def close(self):
        if getattr(self, "h", None):
            self.lib.xpbd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
