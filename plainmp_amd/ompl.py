"""Box / euclidean motion validators (ompl/motion_validator.hpp:42-63).

The reference validators are OMPL MotionValidator subclasses; here they sit
directly on a ``SphereCollisionCst`` and check one edge (``checkMotion``) or
a batch of edges in one persistent launch (``check_motions``).
"""

from __future__ import annotations

from typing import Sequence

from ._native.ompl import EdgeGraph, MotionValidator, ValidatorConfig, ValidatorType  # noqa: F401


def BoxMotionValidator(cst, width: Sequence[float]) -> MotionValidator:
    """BoxMotionValidator(si, width) (motion_validator.cpp:81-111)."""
    vc = ValidatorConfig()
    vc.type = ValidatorType.BOX
    vc.box_width = [float(w) for w in width]
    return MotionValidator(cst, vc)


def EuclideanMotionValidator(cst, resolution: float) -> MotionValidator:
    """EuclideanMotionValidator(si, resolution) (motion_validator.cpp:113-126)."""
    vc = ValidatorConfig()
    vc.type = ValidatorType.EUCLIDEAN
    vc.resolution = float(resolution)
    return MotionValidator(cst, vc)


__all__ = ["ValidatorType", "ValidatorConfig", "MotionValidator", "BoxMotionValidator", "EuclideanMotionValidator",
           "EdgeGraph"]
