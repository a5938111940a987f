"""Algorithmic work model of the FK + sphere-SDF path, for roofline figures.

Convention: every add, sub, mul, abs, min, max, compare and sqrt counts 1
flop and an FMA counts 2; work is full evaluation (no early exit, no
culling).  Per-block costs:

    revolute joint (sincos 25 + local quat 3 + origin product 28 + compose 61)   117
        (89 when the origin rotation is flagged identity)
    prismatic joint                                                               39
    fixed link compose                                                            61 (33 identity)
    rotation matrix per sphere-bearing link                                        24
    sphere world centre R p + t                                                    18
    sphere-primitive pair: box aligned / box rotated / sphere / cylinder z /
        cylinder rotated / ground                                                 25 / 40 / 11 / 23 / 38 / 4
    self sphere pair                                                               11
    adjoint per (sphere, primitive) pair: analytic gradient 12 + 17 per column
    cloud pair as the K = 4 product [x, 1] . [-2y, |y|^2]                           8
    smoothness per waypoint per dof (value + gradient)                             16
    lerp per waypoint per dof                                                       3

MI355X peaks (AMD published): FP32 vector 157.3 TFLOP/s (78.6 without packed
FP32), FP64 vector 78.6, FP32 matrix 157.3, HBM3E 8 TB/s.
"""

from __future__ import annotations

import numpy as np

PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP32_VALU_NONPACKED_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _pair_cost(p) -> int:
    t = p["type"]
    if t == "box":
        return 25 if np.allclose(p["rotation"], np.eye(3)) else 40
    if t == "sphere":
        return 11
    if t == "cylinder":
        R = np.asarray(p["rotation"])
        z_aligned = abs(R[2, 2] - 1) < 1e-6 and abs(R[0, 2]) < 1e-6 and abs(R[1, 2]) < 1e-6
        return 23 if z_aligned else 38
    if t == "ground":
        return 4
    raise KeyError(t)


def fk_flops(robot) -> int:
    """Joint + fixed-link compose + rotation matrices + sphere centres of one configuration."""
    import xml.etree.ElementTree as ET

    root = ET.fromstring(robot.urdf)
    total = 0
    for j in root.findall("joint"):
        jt = j.get("type")
        o = j.find("origin")
        rpy = [float(v) for v in o.get("rpy", "0 0 0").split()] if o is not None else [0, 0, 0]
        identity = all(abs(v) < 2e-6 for v in rpy)
        if jt in ("revolute", "continuous"):
            total += 89 if identity else 117
        elif jt == "prismatic":
            total += 39
        else:
            total += 33 if identity else 61
    links = {s[0] for s in robot.sphere_specs}
    total += 24 * len(links) + 18 * robot.n_spheres
    return total


def config_flops(robot, prims, self_pairs: int = 0) -> int:
    """Full-evaluation flops of one collision check (FK + all pairs)."""
    pair = sum(_pair_cost(p) for p in prims)
    return fk_flops(robot) + robot.n_spheres * pair + 11 * self_pairs


def waypoint_flops(robot, prims, self_pairs: int = 0) -> int:
    """One edge waypoint: lerp of every joint plus a collision check."""
    return 3 * robot.dof + config_flops(robot, prims, self_pairs)


def pair_grad_flops(robot, prims, n_rel_mean: float) -> float:
    """Per configuration: FK + every (sphere, primitive) value and its adjoint."""
    n_pairs = robot.n_spheres * len(prims)
    return fk_flops(robot) + sum(robot.n_spheres * _pair_cost(p) for p in prims) + n_pairs * (12 + 17 * n_rel_mean)
