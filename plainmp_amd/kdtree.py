"""Exact nearest neighbour (plainmp.kdtree; collision/kdtree.hpp:28-47), on the GPU,
and the device build of a cloud's tile tables (the structure the cloud
distance kernels use in place of the KD-tree, collision/kdtree.cpp:40-65):
``cloud_tile_sizes(n) -> (n_tiles, n_supers, scratch_bytes)`` and
``build_cloud_tiles(points (N, 3)) -> {tiles, tiles64, tile_b, points64, supers}``
(the tables as bytes, laid out as pmp_cloud_build_tiles documents)."""

from ._native.kdtree import KDTree, build_cloud_tiles, cloud_tile_sizes  # noqa: F401

__all__ = ["KDTree", "build_cloud_tiles", "cloud_tile_sizes"]
