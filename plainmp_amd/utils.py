"""Conversions from scikit-robot scene objects (src/plainmp/utils.py).

scikit-robot is not a dependency: ``primitive_to_plainmp_sdf`` reads what it
needs from the object -- ``worldpos()``, ``worldrot()`` and ``extents`` /
``radius`` / ``height`` -- so skrobot's ``Box``, ``Sphere`` and ``Cylinder``
(and anything shaped like them) convert without importing skrobot.
"""

from __future__ import annotations

from typing import Sequence, Union

import numpy as np

from . import psdf


def _kind(shape) -> str:
    """skrobot class of the shape (by name along its MRO), else its attributes."""
    names = {c.__name__ for c in type(shape).__mro__}
    for k in ("Sphere", "Box", "Cylinder"):  # the reference's isinstance order (utils.py:44-53)
        if k in names:
            return k
    if hasattr(shape, "extents"):
        return "Box"
    if hasattr(shape, "height") and hasattr(shape, "radius"):
        return "Cylinder"
    if hasattr(shape, "radius"):
        return "Sphere"
    raise ValueError(f"Unsupported shape type {type(shape)}")


def primitive_to_plainmp_sdf(shape) -> psdf.SDFBase:
    """skrobot primitive -> plainmp SDF with the shape's world pose
    (primitive_to_plainmp_sdf, src/plainmp/utils.py:25-57): Sphere -> SphereSDF(radius),
    Box -> BoxSDF(extents), Cylinder -> CylinderSDF(radius, height); anything
    else raises ValueError."""
    kind = _kind(shape)
    pose = psdf.Pose(np.asarray(shape.worldpos(), float), np.asarray(shape.worldrot(), float))
    if kind == "Sphere":
        return psdf.SphereSDF(float(shape.radius), pose)
    if kind == "Box":
        return psdf.BoxSDF(np.asarray(shape.extents, float), pose)
    return psdf.CylinderSDF(float(shape.radius), float(shape.height), pose)


def box_to_grid_poitns(box, N_points: Union[int, Sequence[int]]) -> np.ndarray:
    """World coordinates of an N_x x N_y x N_z grid spanning a box
    (box_to_grid_poitns, src/plainmp/utils.py:13-22; the name is the
    reference's).  Local points in the order of numpy.meshgrid(x, y, z).T,
    mapped by the box's world pose."""
    if isinstance(N_points, int):
        N_points = [N_points, N_points, N_points]
    x, y, z = np.asarray(box.extents, float) * 0.5
    lin = [np.linspace(-x, x, N_points[0]), np.linspace(-y, y, N_points[1]), np.linspace(-z, z, N_points[2])]
    local = np.array(np.meshgrid(*lin)).T.reshape(-1, 3)
    if hasattr(box, "transform_vector"):
        return np.asarray(box.transform_vector(local))
    R = np.asarray(box.worldrot(), float)
    return local @ R.T + np.asarray(box.worldpos(), float)
