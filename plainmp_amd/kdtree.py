"""Exact nearest neighbour (plainmp.kdtree; collision/kdtree.hpp:28-47), on the GPU."""

from ._native.kdtree import KDTree  # noqa: F401

__all__ = ["KDTree"]
