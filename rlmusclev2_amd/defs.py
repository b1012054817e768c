"""Constraint type tags and FEM flags.

Same integer values as the reference ``ConstraintType`` / ``FemFlag``
(src/muscle/defs.py:20-42) so constraint tables and JSON configs carry over.
"""
from __future__ import annotations

from enum import IntEnum, IntFlag


class ConstraintType(IntEnum):
    PIN = 157323
    ATTACH = 1650556
    PINTOTARGET = 771624311
    TETVOLUME = -215389979
    TETFIBERNORM = -303462111
    DISTANCELINE = 1621136047
    TETARAP = -92199131
    TETARAPSVD = -92199132
    TETARAPNORM = -885573303
    TETARAPANISO = -92199133
    TETFIBERDGF = -403562211
    TETSNH = -503462311
    TETFIBERMILLARD = -503662411
    DISTANCE = 1621136100
    TETVOLUMEHARD = -215389990
    TETINVERSION = -215389991


class FemFlag(IntFlag):
    LINEARENERGY = 1 << 0
    NORMSTIFFNESS = 1 << 1


# types the HIP substep implements (the rest raise XpbdError(EUNSUPPORTED))
HIP_TYPES = frozenset({
    ConstraintType.TETARAPSVD, ConstraintType.TETVOLUME, ConstraintType.TETSNH,
    ConstraintType.TETARAPANISO, ConstraintType.ATTACH, ConstraintType.PINTOTARGET,
})
