"""MI355X-native XPBD muscle / soft-tissue substep (drop-in for the substep of
chunleili/rlmuscleV2's XpbdEngine under MuscleSim / MultiGridSim).

Python host code drives hand-written HIP kernels (gfx950) through the C ABI in
include/xpbd.h, loaded with ctypes from rlmusclev2_amd/_lib/libxpbd_hip.so.
"""
from .defs import ConstraintType, FemFlag  # noqa: F401
from .config import load_config, find_constraint  # noqa: F401

__all__ = ["ConstraintType", "FemFlag", "load_config", "find_constraint"]
