"""Primitive signed-distance functions (plainmp.psdf; collision/primitive_sdf.hpp).

evaluate / evaluate_batch / is_outside run on the GPU in FP64.  Points for
``evaluate_batch`` are (3, N), as in the reference (psdf.pyi:58-64).
"""

from ._native.primitive_sdf import (  # noqa: F401
    BoxSDF,
    CloudSDF,
    CylinderSDF,
    GroundSDF,
    Pose,
    PrimitiveSDFBase,
    SDFBase,
    SphereSDF,
    UnionSDF,
)

__all__ = ["Pose", "SDFBase", "PrimitiveSDFBase", "UnionSDF", "GroundSDF", "BoxSDF", "CylinderSDF", "SphereSDF",
           "CloudSDF"]
