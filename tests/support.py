"""Builders shared by the tests: turn a synthetic workload into oracle
objects (and, in GPU tests, product objects) from the same description."""

from __future__ import annotations

import numpy as np

from oracle import _oracle, urdf_parse


def oracle_prim(p):
    t = p["type"]
    if t == "box":
        return _oracle.Prim.box(np.asarray(p["width"], float), np.asarray(p["position"], float),
                                np.asarray(p["rotation"], float))
    if t == "sphere":
        return _oracle.Prim.sphere(float(p["radius"]), np.asarray(p["position"], float),
                                   np.asarray(p["rotation"], float))
    if t == "cylinder":
        return _oracle.Prim.cylinder(float(p["radius"]), float(p["height"]), np.asarray(p["position"], float),
                                     np.asarray(p["rotation"], float))
    if t == "ground":
        return _oracle.Prim.ground(float(p["height"]))
    raise KeyError(t)


def oracle_cst(wl, reorder=True, base_type=None, fixed_prims=()):
    if base_type is None:
        base_type = _oracle.BaseType.FIXED
    kin = urdf_parse.kinematic_model(wl.robot.urdf)
    specs = [_oracle.SphereSpec(ln, np.asarray(pos, float), list(map(float, r)), bool(only))
             for ln, pos, r, only in wl.robot.sphere_specs]
    cst = _oracle.SphereCollisionCst(kin, wl.robot.joint_names, base_type, specs, list(wl.robot.self_pairs),
                                     [oracle_prim(p) for p in fixed_prims], reorder)
    prims = [oracle_prim(p) for p in wl.prims]
    if wl.cloud is not None:
        prims.append(_oracle.Prim.cloud(np.asarray(wl.cloud, float), float(wl.cloud_radius)))
    if prims:
        cst.set_sdf(prims)
    return cst


# ------------------------------------------------------------- product side
def product_sdf(p):
    from plainmp_amd.synth import build_sdf

    return build_sdf(p)


def product_cst(wl, reorder=True, base_type=None, fixed_prims=()):
    from plainmp_amd.synth import build_constraint

    return build_constraint(wl, reorder, base_type, fixed_prims)


def oracle_min_margin(ocst, Q):
    """Per configuration, the smallest |signed clearance| over every sphere /
    primitive pair and self pair (Taylor FK), i.e. how close the boolean is
    to flipping."""
    d = ocst.sphere_clearance_batch(Q, False, 8)
    return np.abs(d).min(axis=1) if d.size else np.full(len(Q), np.inf)


def rows_agree(oc, X, got, ref, sel, row, tol=1e-4, h=1e-6):
    """Evaluate rows of the selected waypoints against the reference's rows
    (relative error <= tol, scale max(|ref|, 1)).  The reference takes the SDF
    gradient by a forward difference of step 1e-6 (primitive_sphere_collision.cpp:308-315):
    where that step straddles a kink of the SDF (inside a box or cylinder,
    near its medial surface) the reference row is not the derivative.  A row
    that differs from the reference's is accepted only if it matches the FP64
    central difference of the oracle's value instead.  Returns (rows checked,
    rows settled by the central difference)."""
    idx = np.flatnonzero(sel)
    g, r = got[idx, row], ref[idx, row]
    bad = ~(np.abs(g - r) <= tol * np.maximum(np.abs(r), 1.0)).all(axis=1)
    if bad.any():
        Xb = X[idx[bad]]
        D = X.shape[1]
        fd = np.empty((len(Xb), D))
        for j in range(D):
            e = np.zeros(D)
            e[j] = h
            fd[:, j] = (oc.evaluate_batch(Xb + e, 8)[0][:, row] - oc.evaluate_batch(Xb - e, 8)[0][:, row]) / (2 * h)
        ok = (np.abs(g[bad] - fd) <= tol * np.maximum(np.abs(fd), 1.0)).all(axis=1)
        assert ok.all(), f"{int((~ok).sum())} rows differ from both the reference row and the central difference"
    return len(idx), int(bad.sum())
