"""Robot specifications from plainmp's YAML configuration files
(src/plainmp/robot_spec.py:72-492, conf/*.yaml), ingested into the device
constraint.

A configuration names the robot's URDF (``urdf_path``), its controlled
joints (``control_joint_names``) and its collision model:

* ``collision_spheres``: ``{link: {"spheres": [[x, y, z, r], ...]}}``, one
  sphere group per link (``parse_sphere_specs``, robot_spec.py:345-380);
* ``only_self_collision_links``: links whose spheres are checked for self
  collision only (fixed base only; a moving base checks every sphere against
  the environment, robot_spec.py:359-365);
* ``self_collision_pairs`` (link pairs) and ``self_collision_group_pairs``
  (every link of ``group1`` with every link of ``group2``), kept in the given
  order, a duplicated pair being an error (robot_spec.py:386-409);
* ``self_body_collision_primitives``: boxes (``extents``), spheres
  (``radius``) and cylinders (``radius``, ``height``) fixed to the base at
  ``position`` with ``rotation`` = [roll, pitch, yaw] (R = Rz(yaw) Ry(pitch)
  Rx(roll)), used as the constraint's fixed SDF (robot_spec.py:263-299 and
  utils.py:25-57; fixed base only);
* ``include``: further YAML files, relative to this one, merged in (their keys
  must not clash; no nested includes; robot_spec.py:100-114).

The URDF is read from ``urdf_path`` (``~`` expanded) unless the XML is given
directly; this module fetches nothing.
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Union

import numpy as np

from .kinematics import BaseType, KinematicModel


def load_config(conf: Union[str, Path, Dict]) -> Dict:
    """The configuration dict of a YAML file (with its includes merged) or a
    dict given directly."""
    import yaml

    if isinstance(conf, dict):
        return dict(conf)
    path = Path(conf)
    with open(path) as f:
        d = yaml.safe_load(f)
    if "include" in d:
        for rel in d["include"]:
            with open(path.parent / rel) as f:
                dep = yaml.safe_load(f)
            if set(d) & set(dep):
                raise ValueError(f"{rel}: keys {sorted(set(d) & set(dep))} are already defined")
            if "include" in dep:
                raise ValueError(f"{rel}: nested include is not supported")
            d.update(dep)
    return d


def rpy_matrix(yaw: float, pitch: float, roll: float) -> np.ndarray:
    """Rz(yaw) Ry(pitch) Rx(roll)."""
    cy, sy = np.cos(yaw), np.sin(yaw)
    cp, sp = np.cos(pitch), np.sin(pitch)
    cr, sr = np.cos(roll), np.sin(roll)
    rz = np.array([[cy, -sy, 0.0], [sy, cy, 0.0], [0.0, 0.0, 1.0]])
    ry = np.array([[cp, 0.0, sp], [0.0, 1.0, 0.0], [-sp, 0.0, cp]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cr, -sr], [0.0, sr, cr]])
    return rz @ ry @ rx


class RobotSpec:
    """A robot described by a plainmp YAML configuration (RobotSpec,
    robot_spec.py:72-160), with the collision constraint built on this
    library's device path."""

    def __init__(self, conf: Union[str, Path, Dict], base_type: BaseType = BaseType.FIXED,
                 urdf: Optional[str] = None):
        self.conf_dict = load_config(conf)
        self.base_type = base_type
        self._urdf = urdf
        self._kin: Optional[KinematicModel] = None

    @property
    def control_joint_names(self) -> List[str]:
        return list(self.conf_dict["control_joint_names"])

    def urdf_string(self) -> str:
        if self._urdf is not None:
            return self._urdf
        path = os.path.expanduser(self.conf_dict["urdf_path"])
        with open(path) as f:
            return f.read()

    def get_kin(self) -> KinematicModel:
        if self._kin is None:
            self._kin = KinematicModel(self.urdf_string())
        return self._kin

    def parse_sphere_specs(self):
        """One SphereAttachmentSpec per link of ``collision_spheres``
        (positions as 3 x k, radii k), in the configuration's order."""
        from .constraint import SphereAttachmentSpec

        only_self = set(self.conf_dict.get("only_self_collision_links", []))
        specs = []
        for link, vals in self.conf_dict["collision_spheres"].items():
            rows = np.array(vals["spheres"], dtype=float).reshape(-1, 4)
            flag = link in only_self if self.base_type == BaseType.FIXED else False
            specs.append(SphereAttachmentSpec(link, np.ascontiguousarray(rows[:, :3].T), rows[:, 3].copy(), flag))
        names = [s.parent_link_name for s in specs]
        if len(set(names)) != len(names):
            raise ValueError("duplicated parent link name")
        return specs

    def get_self_collision_pairs(self) -> List[List[str]]:
        pairs = [list(p) for p in self.conf_dict.get("self_collision_pairs", [])]
        for g in self.conf_dict.get("self_collision_group_pairs", []):
            pairs.extend([a, b] for a in g["group1"] for b in g["group2"])
        seen = set()
        for p in pairs:
            key = tuple(sorted(p))
            if key in seen:
                raise ValueError(f"Duplicated self collision pair: {key}")
            seen.add(key)
        return pairs

    def self_body_collision_primitives(self) -> list:
        """The base-fixed primitives as SDFs (BoxSDF with the full extents,
        SphereSDF, CylinderSDF), posed at ``position`` with the rotation of
        ``rotation`` = [roll, pitch, yaw]."""
        from . import psdf

        out = []
        for d in self.conf_dict.get("self_body_collision_primitives", []):
            roll, pitch, yaw = (float(v) for v in d["rotation"])
            pose = psdf.Pose(np.asarray(d["position"], float), rpy_matrix(yaw, pitch, roll))
            t = d["type"]
            if t == "box":
                out.append(psdf.BoxSDF(np.asarray(d["extents"], float), pose))
            elif t == "sphere":
                out.append(psdf.SphereSDF(float(d["radius"]), pose))
            elif t == "cylinder":
                out.append(psdf.CylinderSDF(float(d["radius"]), float(d["height"]), pose))
            else:
                raise ValueError("Invalid primitive type")
        if self.base_type != BaseType.FIXED and out:
            raise ValueError("self_body_collision_primitives are only used for the fixed base robot; define "
                             "collision spheres instead for a floating or planar base")
        return out

    def create_collision_const(self, self_collision: bool = True, attachments: Sequence = (),
                               reorder_spheres: bool = True):
        """SphereCollisionCst of this robot (create_collision_const,
        robot_spec.py:411-492): the configuration's spheres with the
        attachments merged into their links (or added as new groups), the
        self-collision pairs and the base-fixed primitives when
        ``self_collision`` is set."""
        from . import psdf
        from .constraint import SphereAttachmentSpec, SphereCollisionCst

        specs = self.parse_sphere_specs()
        for att in attachments:
            for i, s in enumerate(specs):
                if s.parent_link_name == att.parent_link_name:
                    specs[i] = SphereAttachmentSpec(
                        s.parent_link_name, np.hstack([s.relative_positions, att.relative_positions]),
                        np.hstack([s.radii, att.radii]), s.only_self_collision)
                    break
            else:
                specs.append(att)
        pairs: List[List[str]] = []
        fixed = None
        if self_collision:
            pairs = self.get_self_collision_pairs()
            prims = self.self_body_collision_primitives()
            if prims:
                fixed = psdf.UnionSDF(prims)
        return SphereCollisionCst(self.get_kin(), self.control_joint_names, self.base_type, specs, pairs, fixed,
                                  reorder_spheres)


__all__ = ["RobotSpec", "load_config", "rpy_matrix"]
