"""skrobot primitive -> SDF conversion (CPU): duck-typed stand-ins shaped
like skrobot's Box / Sphere / Cylinder (worldpos(), worldrot(), extents,
radius, height); the resulting primitives' AABBs equal the oracle's for the
same shape and pose, and unsupported shapes raise ValueError
(src/plainmp/utils.py:25-57)."""

import math

import numpy as np
import pytest

from oracle import _oracle
from plainmp_amd import psdf
from plainmp_amd.utils import box_to_grid_poitns, primitive_to_plainmp_sdf


class _Shape:
    def __init__(self, pos, rot):
        self._p, self._R = np.asarray(pos, float), np.asarray(rot, float)

    def worldpos(self):
        return self._p

    def worldrot(self):
        return self._R


class Box(_Shape):
    def __init__(self, extents, pos, rot):
        super().__init__(pos, rot)
        self.extents = list(extents)


class Sphere(_Shape):
    def __init__(self, radius, pos, rot):
        super().__init__(pos, rot)
        self.radius = radius


class Cylinder(_Shape):
    def __init__(self, radius, height, pos, rot):
        super().__init__(pos, rot)
        self.radius, self.height = radius, height


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def test_primitive_to_plainmp_sdf_matches_oracle_shapes():
    R = _rz(0.4) @ np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    p = [0.7, -0.2, 0.9]
    cases = [
        (Box([1.0, 2.0, 0.05], p, R), psdf.BoxSDF, _oracle.Prim.box(np.array([1.0, 2.0, 0.05]), np.array(p), R)),
        (Sphere(0.3, p, np.eye(3)), psdf.SphereSDF, _oracle.Prim.sphere(0.3, np.array(p), np.eye(3))),
        (Cylinder(0.1, 0.6, p, R), psdf.CylinderSDF, _oracle.Prim.cylinder(0.1, 0.6, np.array(p), R)),
    ]
    for shape, cls, ref in cases:
        sdf = primitive_to_plainmp_sdf(shape)
        assert isinstance(sdf, cls)
        assert np.allclose(sdf.pose.position, p) and np.allclose(sdf.pose.rotation, shape.worldrot())
        assert np.allclose(sdf.lb, ref.lb, atol=1e-12) and np.allclose(sdf.ub, ref.ub, atol=1e-12)


def test_unsupported_shape_raises():
    class Mesh(_Shape):
        pass

    with pytest.raises(ValueError, match="Unsupported shape type"):
        primitive_to_plainmp_sdf(Mesh([0, 0, 0], np.eye(3)))


def test_box_to_grid_points():
    box = Box([1.0, 2.0, 4.0], [1.0, 0.0, 0.5], _rz(math.pi / 2))
    pts = box_to_grid_poitns(box, [2, 3, 2])
    assert pts.shape == (12, 3)
    local = (pts - [1.0, 0.0, 0.5]) @ _rz(math.pi / 2)  # back to the box frame
    assert np.allclose(np.sort(np.unique(local[:, 0].round(12))), [-0.5, 0.5])
    assert np.allclose(np.sort(np.unique(local[:, 1].round(12))), [-1.0, 0.0, 1.0])
    assert np.allclose(np.sort(np.unique(local[:, 2].round(12))), [-2.0, 2.0])
