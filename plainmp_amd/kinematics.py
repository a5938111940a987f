"""Kinematic model (plainmp.kinematics; kinematics/kinematics.hpp:51-169)."""

from ._native.kinematics import BaseType, KinematicModel  # noqa: F401

__all__ = ["BaseType", "KinematicModel"]
