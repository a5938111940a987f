"""URDF -> oracle KinematicModel, with Python's ElementTree.

Test infrastructure only.  This reader is deliberately independent of the
product's C++ URDF reader so that the two can check each other.  Numerics
follow urdfdom as vendored by the reference:
  * origin rpy -> quaternion, then normalised
    (third/urdf_parser/include/urdf_model/pose.h:202-250);
  * joint axis parsed as written, not normalised, default (1, 0, 0)
    (third/urdf_parser/src/joint.cpp:427-452);
  * limit lower/upper default to 0 (joint.cpp:101-130);
  * children ordered by joint name (std::map in model.h:113-153).
"""

from __future__ import annotations

import xml.etree.ElementTree as ET

from . import _oracle

JOINT_TYPES = {
    "revolute": 1,
    "continuous": 2,
    "prismatic": 3,
    "floating": 4,
    "planar": 5,
    "fixed": 6,
}


def _vec(text, default):
    if text is None:
        return list(default)
    vals = [float(v) for v in text.split()]
    if len(vals) != 3:
        raise ValueError(f"malformed vector {text!r}")
    return vals


def parse(xml: str):
    """Return (link_names, joint_tuples) for ``_oracle.KinematicModel``."""
    if not xml:
        raise RuntimeError("xml string is empty")
    root = ET.fromstring(xml)
    links = [ln.get("name") for ln in root.findall("link")]
    joints = []
    for j in root.findall("joint"):
        jtype = JOINT_TYPES[j.get("type")]
        origin = j.find("origin")
        xyz = _vec(origin.get("xyz") if origin is not None else None, (0, 0, 0))
        rpy = _vec(origin.get("rpy") if origin is not None else None, (0, 0, 0))
        quat = _oracle.urdf_quat_from_rpy(*rpy)
        axis_el = j.find("axis")
        axis = [1.0, 0.0, 0.0]
        if axis_el is not None and axis_el.get("xyz") is not None:
            axis = _vec(axis_el.get("xyz"), (1, 0, 0))
        lim = j.find("limit")
        lower = float(lim.get("lower", 0.0)) if lim is not None else 0.0
        upper = float(lim.get("upper", 0.0)) if lim is not None else 0.0
        joints.append(
            (
                j.get("name"),
                jtype,
                j.find("parent").get("link"),
                j.find("child").get("link"),
                xyz,
                quat,
                axis,
                lower,
                upper,
            )
        )
    return links, joints


def kinematic_model(xml: str) -> "_oracle.KinematicModel":
    links, joints = parse(xml)
    return _oracle.KinematicModel(links, joints)
