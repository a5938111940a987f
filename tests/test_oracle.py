"""Pins the FP64 oracle (CPU only).

Each check is independent of the oracle's own formulas: the reference's
known answers and invariants (cited), an independent numpy FK with 4x4
homogeneous matrices and Rodrigues rotations, closed-form SDFs written
differently, central finite differences and brute force.
"""

import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import _oracle, urdf_parse
from plainmp_amd import synth
from support import oracle_cst, oracle_prim

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- Taylor
def test_taylor_sincos_bounds_and_special_cases():
    s, c = _oracle.taylor_sincos(0.0)
    assert s == 0.0 and c == 1.0
    xs = np.linspace(-1.45, 1.45, 2001)
    err_s = max(abs(_oracle.taylor_sincos(x)[0] - math.sin(x)) for x in xs)
    err_c = max(abs(_oracle.taylor_sincos(x)[1] - math.cos(x)) for x in xs)
    # degree-9 / degree-8 remainders at |x| = 1.45: |x|^11/11! and |x|^10/10!
    assert err_s <= 1.45 ** 11 / math.factorial(11) * 1.01
    assert err_c <= 1.45 ** 10 / math.factorial(10) * 1.01
    assert err_s > 1e-9  # a truncated series, not the exact function
    # range reduction: periodicity and the +-pi/2 folds
    for x in (-7.0, -3.5, -2.0, 2.0, 3.5, 7.0, 100.0):
        s, c = _oracle.taylor_sincos(x)
        assert abs(s - math.sin(x)) < 4e-6 and abs(c - math.cos(x)) < 3e-5
    for x in (math.pi / 2, -math.pi / 2):
        s, c = _oracle.taylor_sincos(x)
        assert abs(s - math.sin(x)) < 4e-6 and abs(c) < 3e-5


# -------------------------------------------------------------------- FK
def _rpy_matrix(r, p, y):
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _rodrigues(axis, angle):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def numpy_fk(urdf, q_by_joint):
    """World 4x4 pose of every link from the URDF text, independent of the oracle."""
    root = ET.fromstring(urdf)
    joints = {}
    for j in root.findall("joint"):
        o = j.find("origin")
        xyz = [float(v) for v in o.get("xyz").split()]
        rpy = [float(v) for v in o.get("rpy").split()]
        ax = j.find("axis")
        axis = [float(v) for v in ax.get("xyz").split()] if ax is not None else [1, 0, 0]
        joints[j.find("child").get("link")] = (j.get("name"), j.get("type"), j.find("parent").get("link"), xyz, rpy, axis)
    poses = {}

    def pose(link):
        if link in poses:
            return poses[link]
        if link not in joints:
            poses[link] = np.eye(4)
            return poses[link]
        name, jt, parent, xyz, rpy, axis = joints[link]
        T = np.eye(4)
        T[:3, :3] = _rpy_matrix(*rpy)
        T[:3, 3] = xyz
        J = np.eye(4)
        if jt in ("revolute", "continuous"):
            J[:3, :3] = _rodrigues(axis, q_by_joint.get(name, 0.0))
        elif jt == "prismatic":
            J[:3, 3] = np.asarray(axis) * q_by_joint.get(name, 0.0)
        poses[link] = pose(parent) @ T @ J
        return poses[link]

    for ln in [ln.get("name") for ln in root.findall("link")]:
        pose(ln)
    return poses


def _quat_to_matrix(qx, qy, qz, qw):
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
    ])


@pytest.mark.parametrize("robot_fn", [synth.panda_like, synth.mobile_torso, synth.dual_arm, synth.fetch_like,
                                      synth.humanoid31])
def test_fk_exact_matches_numpy_homogeneous(robot_fn):
    robot = robot_fn(0)
    kin = urdf_parse.kinematic_model(robot.urdf)
    ids = kin.get_joint_ids(robot.joint_names)
    links, _ = urdf_parse.parse(robot.urdf)
    for _ in range(10):
        q = RNG.uniform(robot.lower, robot.upper)
        kin.set_joint_positions(ids, list(q), True)
        ref = numpy_fk(robot.urdf, dict(zip(robot.joint_names, q)))
        for ln in links:
            p = kin.get_link_pose(kin.get_link_ids([ln])[0])
            assert np.allclose(p[:3], ref[ln][:3, 3], atol=1e-12)
            assert np.allclose(_quat_to_matrix(*p[3:]), ref[ln][:3, :3], atol=1e-12)


def test_quat_to_rpy_round_trip_and_gimbal_lock():
    """getRPY (transform.hpp:93-120): inverse of setQuaternionFromRPY away from
    |pitch| = pi/2; the two gimbal-lock branches as known answers."""
    rng = np.random.default_rng(3)

    def rot(q):
        x, y, z, w = q
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])

    for _ in range(200):
        r, p, y = rng.uniform(-3.1, 3.1), rng.uniform(-1.5, 1.5), rng.uniform(-3.1, 3.1)
        assert np.allclose(_oracle.quat_to_rpy(*_oracle.quat_from_rpy(r, p, y)), [r, p, y], atol=1e-9)
    # at the lock only yaw -/+ roll is defined.  The reference's branch for
    # pitch = -pi/2 (yaw = -2 atan2(x, y)) reproduces the rotation; the one
    # for +pi/2 (yaw = 2 atan2(x, y)) returns roll - yaw where yaw - roll
    # would, a quirk kept as the reference computes it
    wrap = lambda a: (a + math.pi) % (2 * math.pi) - math.pi  # noqa: E731
    for r, y in [(0.4, -0.7), (0.0, 0.5), (-0.2, 1.0)]:
        q = _oracle.quat_from_rpy(r, -math.pi / 2, y)
        rr = _oracle.quat_to_rpy(*q)
        assert rr[0] == 0.0 and rr[1] == -math.pi / 2
        assert np.allclose(rot(_oracle.quat_from_rpy(*rr)), rot(q), atol=1e-9)
        rr = _oracle.quat_to_rpy(*_oracle.quat_from_rpy(r, math.pi / 2, y))
        assert rr[0] == 0.0 and rr[1] == math.pi / 2
        assert abs(wrap(rr[2] - (r - y))) < 1e-9


def test_fk_known_answers():
    robot = synth.panda_like(0)
    kin = urdf_parse.kinematic_model(robot.urdf)
    ids = kin.get_joint_ids(robot.joint_names)
    kin.set_joint_positions(ids, [0.0] * 7, True)
    ref = numpy_fk(robot.urdf, {})
    p1 = kin.get_link_pose(kin.get_link_ids(["link1"])[0])
    root = ET.fromstring(robot.urdf)
    j1 = [j for j in root.findall("joint") if j.get("name") == "joint1"][0]
    assert np.allclose(p1[:3], [float(v) for v in j1.find("origin").get("xyz").split()])
    # a single revolute joint rotates its child about z by the angle
    kin.set_joint_positions([ids[0]], [0.7], True)
    p = kin.get_link_pose(kin.get_link_ids(["link1"])[0])
    R = _quat_to_matrix(*p[3:])
    assert np.allclose(R, _rodrigues([0, 0, 1], 0.7), atol=1e-12)
    assert np.allclose(ref["link1"][:3, 3], p1[:3])


def test_taylor_fk_close_to_exact():
    robot = synth.panda_like(0)
    kin = urdf_parse.kinematic_model(robot.urdf)
    ids = kin.get_joint_ids(robot.joint_names)
    tip = kin.get_link_ids(["hand"])[0]
    worst = 0.0
    for _ in range(200):
        q = list(RNG.uniform(-2.9, 2.9, size=7))
        kin.set_joint_positions(ids, q, False)
        a = np.array(kin.get_link_pose(tip)[:3])
        kin.set_joint_positions(ids, q, True)
        b = np.array(kin.get_link_pose(tip)[:3])
        worst = max(worst, np.abs(a - b).max())
    assert 1e-8 < worst < 1e-4


def test_point_jacobian_vs_central_fd():
    robot = synth.panda_like(0)
    kin = urdf_parse.kinematic_model(robot.urdf)
    ids = kin.get_joint_ids(robot.joint_names)
    link = kin.get_link_ids(["link6"])[0]
    local = np.array([0.03, -0.02, 0.05])
    h = 1e-6

    def point(q):
        kin.set_joint_positions(ids, list(q), True)
        p = kin.get_link_pose(link)
        return np.array(p[:3]) + _quat_to_matrix(*p[3:]) @ local

    for _ in range(10):
        q = RNG.uniform(-2.9, 2.9, size=7)
        c = point(q)
        J = kin.get_attached_point_jacobian(link, c, ids, _oracle.BaseType.FIXED)
        fd = np.stack([(point(q + h * e) - point(q - h * e)) / (2 * h) for e in np.eye(7)], axis=1)
        assert np.allclose(J, fd, atol=1e-8)
        assert np.all(J[:, 6] == 0.0)  # joint 7 does not move link 6


# ------------------------------------------------------------------ SDFs
def _box_sdf(p, c, R, w):
    local = (p - c) @ R  # R^T (p - c)
    d = np.abs(local) - w / 2
    outside = np.linalg.norm(np.maximum(d, 0), axis=-1)
    inside = np.minimum(d.max(-1), 0)
    return outside + inside


def _cyl_sdf(p, c, R, r, h):
    local = (p - c) @ R
    rho = np.hypot(local[:, 0], local[:, 1])
    a = rho - r
    b = np.abs(local[:, 2]) - h / 2
    return np.hypot(np.maximum(a, 0), np.maximum(b, 0)) + np.minimum(np.maximum(a, b), 0)


def test_sdf_closed_forms_and_is_outside():
    pts = RNG.normal(0.5, 0.7, size=(400, 3))
    for p in synth.mixed_primitives(30, seed=9):
        o = oracle_prim(p)
        got = o.evaluate_batch(pts.T)
        if p["type"] == "box":
            ref = _box_sdf(pts, p["position"], p["rotation"], p["width"])
        elif p["type"] == "sphere":
            ref = np.linalg.norm(pts - p["position"], axis=1) - p["radius"]
        else:
            ref = _cyl_sdf(pts, p["position"], p["rotation"], p["radius"], p["height"])
        assert np.allclose(got, ref, atol=1e-12)
        for r in (0.0, 0.05, 0.3):  # tests/test_psdf.py:165-169
            out = np.array([o.is_outside(x, r) for x in pts])
            assert np.array_equal(out, got > r)


def test_sdf_known_answers():
    # tests/test_psdf.py:89-102
    b = _oracle.Prim.box(np.ones(3), np.zeros(3), np.eye(3))
    b.translate(np.array([1.0, 2.0, 3.0]))
    assert np.allclose(b.lb, [0.5, 1.5, 2.5]) and np.allclose(b.ub, [1.5, 2.5, 3.5])
    b = _oracle.Prim.box(np.ones(3), np.zeros(3), np.eye(3))
    b.rotate_z(math.pi / 4)
    s = 0.5 * math.sqrt(2)
    assert np.allclose(b.lb, [-s, -s, -0.5]) and np.allclose(b.ub, [s, s, 0.5])
    # tests/test_psdf.py:27-40: alignment flags
    assert _oracle.Prim.box(np.ones(3), np.zeros(3), np.eye(3)).axis_aligned
    yaw = _rpy_matrix(0.0, 0.0, 0.3)
    p2 = _oracle.Prim.box(np.ones(3), np.zeros(3), yaw)
    assert not p2.axis_aligned and p2.z_axis_aligned
    p3 = _oracle.Prim.box(np.ones(3), np.zeros(3), _rpy_matrix(0.3, 0.3, 0.3))
    assert not p3.axis_aligned and not p3.z_axis_aligned
    # tests/test_psdf.py:51-57: rotate_z keeps the position
    p4 = _oracle.Prim.box(np.ones(3), np.array([1.0, 2.0, 3.0]), np.eye(3))
    p4.rotate_z(0.3)
    assert np.allclose(p4.position, [1, 2, 3]) and np.allclose(p4.rotation, yaw)
    # tests/test_psdf.py:201-207: ground
    g = _oracle.Prim.ground(1.0)
    assert g.evaluate(np.zeros(3)) == 1.0 and g.evaluate(np.array([1.0, 0, 0])) == 1.0
    assert np.array_equal(g.lb[2:], [0.0]) and np.array_equal(g.ub[2:], [0.0])


def test_kdtree_vs_brute_force():
    # tests/test_kdtree.py:6-20
    pts = RNG.normal(size=(1000, 3))
    tree = _oracle.KDTree(pts)
    for _ in range(300):
        q = RNG.normal(size=3)
        i = np.argmin(((pts - q) ** 2).sum(1))
        assert np.allclose(tree.query(q), pts[i])
        assert np.isclose(tree.sqdist(q), ((pts[i] - q) ** 2).sum())
    cloud = _oracle.Prim.cloud(pts, 0.1)
    q = RNG.normal(size=(3, 50))
    brute = np.sqrt(((pts[None, :, :] - q.T[:, None, :]) ** 2).sum(-1).min(1)) - 0.1
    assert np.allclose(cloud.evaluate_batch(q), brute, atol=1e-12)


# ----------------------------------------------------------- collision cst
def _morton_codes(P):
    """30-bit Morton codes of points in their bounds, interleaved bit by bit
    (x in bit 3 b, y in 3 b + 1, z in 3 b + 2 for bit b of each 10-bit
    coordinate)."""
    lo, hi = P.min(0), P.max(0)
    span = hi - lo
    u = np.where(span > 0, (P - lo) / np.where(span > 0, span, 1.0), 0.0)
    c = np.clip(u * 1023.0, 0.0, 1023.0).astype(np.uint64)
    code = np.zeros(len(P), np.uint64)
    for b in range(10):
        for k in range(3):
            code |= ((c[:, k] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + k)
    return code


@pytest.mark.parametrize("n", [1, 64, 65, 1000, 5000])
def test_cloud_tiles_invariants(n):
    """The oracle's tile tables: the tiles hold the points in (Morton code,
    index) order, exactly once; FP64 bounds are the min / max of each tile's
    points and the FP32 bounds enclose them; the B operands give
    |c - p|^2 = |c - o|^2 + |p'|^2 - 2 (c - o).p'; padding is far away;
    super-tiles cover 16 tiles under the union of their bounds."""
    rng = np.random.default_rng(n)
    P = rng.uniform(-1, 1, (n, 3)) * np.array([2.0, 0.5, 1.0])
    d = _oracle.cloud_tiles(P)
    t32 = np.frombuffer(d["tiles"], np.float32).reshape(-1, 12)
    tn = np.frombuffer(d["tiles"], np.int32).reshape(-1, 12)[:, 9]
    t64 = np.frombuffer(d["tiles64"], np.float64).reshape(-1, 6)
    tb = np.frombuffer(d["tile_b"], np.float32).reshape(-1, 4, 4, 16)  # [tile][m][k][j]
    p64 = np.frombuffer(d["points64"], np.float64).reshape(-1, 64, 4)
    sup = np.frombuffer(d["supers"], np.float32).reshape(-1, 8)
    supi = np.frombuffer(d["supers"], np.int32).reshape(-1, 8)
    nt = (n + 63) // 64
    assert len(t32) == nt and len(p64) == nt and len(tb) == nt and len(sup) == (nt + 15) // 16
    order = np.lexsort((np.arange(n), _morton_codes(P)))
    live = p64[:, :, :3].reshape(-1, 3)[:n]
    assert np.array_equal(live, P[order])
    assert np.all(p64.reshape(-1, 4)[n:, :3] == 1e30)
    for t in range(nt):
        pts = P[order[64 * t:64 * t + 64]]
        assert tn[t] == len(pts)
        assert np.array_equal(t64[t, :3], pts.min(0)) and np.array_equal(t64[t, 3:], pts.max(0))
        assert np.all(t32[t, 0:3].astype(np.float64) < pts.min(0)) and np.all(t32[t, 3:6].astype(np.float64) > pts.max(0))
        o = t32[t, 6:9].astype(np.float64)
        assert np.allclose(o, 0.5 * (pts.min(0) + pts.max(0)), atol=1e-7)
        B = tb[t].transpose(0, 2, 1).reshape(64, 4).astype(np.float64)  # point 16 m + j -> (-2p', |p'|^2)
        c = rng.uniform(-1, 1, (5, 3))
        d2 = ((c - o) ** 2).sum(1)[:, None] + B[None, :len(pts), 3] + (c - o) @ B[:len(pts), :3].T
        ref = ((c[:, None, :] - pts[None]) ** 2).sum(2)
        assert np.abs(d2 - ref).max() <= 1e-6
        assert np.all(B[len(pts):, :3] == 0) and np.all(B[len(pts):, 3] == np.float32(1e30))
    for s in range(len(sup)):
        first, count = supi[s, 6], supi[s, 7]
        assert first == 16 * s and count == min(16, nt - 16 * s)
        assert np.array_equal(sup[s, 0:3], t32[first:first + count, 0:3].min(0))
        assert np.array_equal(sup[s, 3:6], t32[first:first + count, 3:6].max(0))


@pytest.mark.parametrize("extent", [0.5, 2.0, 4.0])
def test_cloud_tile_local_fp32_error_bound(extent):
    """The FP32 tile-local squared distance |c'|^2 + (|p'|^2 - 2 c'.p') from
    the tile tables, emulated with every product and sum rounded to single
    precision (the matrix-core K = 4 order) and as an FMA chain (the edge
    pass's order), stays within 8u (|c'| + rho)^2 of the exact |c - p|^2 and so
    within half the kernels' band term 2e-6 (|c'|^2 + rho^2).  For tiles
    spread over metres the error exceeds the 1e-5 m validity band
    (1e-5 (2R + 1e-5) ~ 6.4e-7 m^2 at R = 0.032), which is why the decisions
    carry the per-tile term."""
    rng = np.random.default_rng(int(extent * 10))
    P = rng.uniform(-extent / 2, extent / 2, (1000, 3)) + np.array([0.3, -0.2, 0.7])
    d = _oracle.cloud_tiles(P)
    t32 = np.frombuffer(d["tiles"], np.float32).reshape(-1, 12)
    tb = np.frombuffer(d["tile_b"], np.float32).reshape(-1, 4, 4, 16)
    p64 = np.frombuffer(d["points64"], np.float64).reshape(-1, 64, 4)
    f32, u = np.float32, 2.0 ** -24
    worst, worst_ratio = 0.0, 0.0
    for t in range(len(t32)):
        lb, ub, o = t32[t, 0:3], t32[t, 3:6], t32[t, 6:9]
        n = int(np.frombuffer(t32[t].tobytes(), np.int32)[9])
        B = tb[t].transpose(0, 2, 1).reshape(64, 4)[:n]  # (-2p', |p'|^2) per point, float32
        pts = p64[t, :n, :3]
        e = np.maximum(ub - o, o - lb).astype(f32)
        rho2 = float(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
        # queries: near the tile's points (the decisions that matter) and anywhere in the cloud
        near = pts[rng.integers(0, n, 40)] + rng.normal(0, 0.05, (40, 3))
        far = rng.uniform(P.min(0), P.max(0), (20, 3))
        C = np.concatenate([near, far]).astype(f32)
        cl = (C - o).astype(f32)  # c' rounded once, as in the kernels
        off = (cl[:, 0] * cl[:, 0] + cl[:, 1] * cl[:, 1] + cl[:, 2] * cl[:, 2]).astype(f32)
        # rounded K = 4 dot product in matrix-core order
        m = (cl[:, None, 0] * B[None, :, 0]).astype(f32)
        for k in (1, 2):
            m = (m + (cl[:, None, k] * B[None, :, k]).astype(f32)).astype(f32)
        m = (m + B[None, :, 3]).astype(f32)
        d_mfma = (m + off[:, None]).astype(f32).astype(np.float64)
        # FMA chain fma(x, c'x, fma(y, c'y, fma(z, c'z, w))): products exact in FP64, one rounding per step
        acc = B[None, :, 3].astype(np.float64)
        for k in (2, 1, 0):
            acc = (B[None, :, k].astype(np.float64) * cl[:, None, k].astype(np.float64) + acc).astype(f32).astype(np.float64)
        d_fma = (acc + off[:, None].astype(np.float64)).astype(f32).astype(np.float64)
        exact = ((C.astype(np.float64)[:, None, :] - pts[None]) ** 2).sum(2)
        a = np.sqrt(off.astype(np.float64))[:, None]
        bound8 = 8 * u * (a + math.sqrt(rho2)) ** 2
        band_term = 2e-6 * (off.astype(np.float64)[:, None] + rho2)
        for dd in (d_mfma, d_fma):
            err = np.abs(dd - exact)
            assert np.all(err <= bound8)
            assert np.all(err <= 0.5 * band_term)
            worst = max(worst, err.max())
            worst_ratio = max(worst_ratio, (err / band_term).max())
    if extent >= 4.0:
        assert worst > 1e-5 * (2 * 0.032 + 1e-5)  # beyond the plain band: the per-tile term is needed
    assert worst_ratio < 0.5


def test_is_valid_consistent_with_evaluate():
    # tests/test_constraints.py:47-53
    wl = synth.workload("grad_mixed32", scale=1 / 64)
    cst = oracle_cst(wl)
    for q in wl.Q[:100]:
        vals, _ = cst.evaluate(q)
        assert bool(np.all(vals > 0)) == cst.is_valid(q)


def test_collision_logic_vs_brute_force_on_centres():
    wl = synth.workload("grad_mixed32", scale=1 / 64)
    cst = oracle_cst(wl)
    prims = [oracle_prim(p) for p in wl.prims]
    C = cst.sphere_centers_batch(wl.Q, False)
    _, radii = cst.get_all_spheres()
    for b, q in enumerate(wl.Q):
        d = np.array([[p.evaluate(C[b, s]) - radii[s] for p in prims] for s in range(len(radii))])
        assert bool((d > 0).all()) == cst.is_valid(q)


def test_evaluate_rows_vs_forward_fd():
    # tests/test_constraints.py:22-44 (forward differences, eps 1e-7, decimal 4)
    wl = synth.workload("grad_mixed32", scale=1 / 256)
    cst = oracle_cst(wl)
    eps = 1e-7
    for q in wl.Q[:20]:
        v0, J = cst.evaluate(q)
        fd = np.stack([(cst.evaluate(q + eps * e)[0] - v0) / eps for e in np.eye(len(q))], axis=1)
        np.testing.assert_almost_equal(J, fd, decimal=4)


def test_empty_constraint_always_valid():
    # tests/test_constraints.py:163-175
    robot = synth.panda_like(0)
    wl = synth.Workload("empty", robot, [])
    cst = oracle_cst(wl)
    assert cst.cst_dim() == 0
    for q in RNG.uniform(-2.9, 2.9, size=(20, 7)):
        assert cst.is_valid(q)
        v, J = cst.evaluate(q)
        assert v.shape == (0,) and J.shape == (0, 7)


def _one_link_robot():
    urdf = """<robot name="r"><link name="base"/><link name="l1"/>
    <joint name="j1" type="revolute"><parent link="base"/><child link="l1"/>
    <origin xyz="0 0 0" rpy="0 0 0"/><axis xyz="0 0 1"/><limit lower="-3" upper="3" effort="1" velocity="1"/></joint>
    </robot>"""
    return urdf


def test_boundary_conventions_ext_and_self():
    # ext: touching collides (strict is_outside); self: touching is free (strict <)
    urdf = _one_link_robot()
    kin = urdf_parse.kinematic_model(urdf)
    spec = _oracle.SphereSpec("l1", np.array([[1.0, 0.0, 0.0]]), [0.25], False)
    box = _oracle.Prim.box(np.array([0.5, 0.5, 0.5]), np.array([1.5, 0.0, 0.0]), np.eye(3))
    cst = _oracle.SphereCollisionCst(kin, ["j1"], _oracle.BaseType.FIXED, [spec], [], [], False)
    cst.set_sdf([box])  # sphere surface at x = 1.25 touches the box face x = 1.25
    assert not cst.is_valid(np.array([0.0]))
    spec_a = _oracle.SphereSpec("l1", np.array([[0.0, 0.0, 0.0]]), [0.25], False)
    spec_b = _oracle.SphereSpec("base", np.array([[0.5, 0.0, 0.0]]), [0.25], False)
    cst2 = _oracle.SphereCollisionCst(urdf_parse.kinematic_model(urdf), ["j1"], _oracle.BaseType.FIXED,
                                      [spec_a, spec_b], [("l1", "base")], [], False)
    assert cst2.is_valid(np.array([0.0]))  # |c1 - c2| = 0.5 = r1 + r2


def test_ground_aabb_quirk():
    # GroundSDF AABB is z in [0, 0] (primitive_sdf.hpp:250-255): a sphere
    # entirely below -r is culled, so a robot under the ground reads valid.
    urdf = _one_link_robot()
    kin = urdf_parse.kinematic_model(urdf)
    ground = _oracle.Prim.ground(0.0)
    for z, expected in ((-1.0, True), (0.2, False), (1.0, True)):
        spec = _oracle.SphereSpec("l1", np.array([[0.0, 0.0, z]]), [0.25], False)
        cst = _oracle.SphereCollisionCst(kin, ["j1"], _oracle.BaseType.FIXED, [spec], [], [], False)
        cst.set_sdf([ground])
        assert cst.is_valid(np.array([0.0])) == expected


def test_prismatic_ignores_origin_rotation():
    # set_joint_angles_impl: prismatic child quaternion is identity (algorithm.cpp:151-156)
    urdf = """<robot name="r"><link name="base"/><link name="l1"/>
    <joint name="j1" type="prismatic"><parent link="base"/><child link="l1"/>
    <origin xyz="0 0 1" rpy="0.5 0 0"/><axis xyz="0 0 1"/><limit lower="0" upper="1" effort="1" velocity="1"/></joint>
    </robot>"""
    kin = urdf_parse.kinematic_model(urdf)
    ids = kin.get_joint_ids(["j1"])
    kin.set_joint_positions(ids, [0.3], True)
    p = kin.get_link_pose(kin.get_link_ids(["l1"])[0])
    assert np.allclose(p[:3], [0, 0, 1.3]) and np.allclose(p[3:], [0, 0, 0, 1])


def test_identity_flag_drops_tiny_rotations():
    # |w - 1| < 1e-6 marks an origin rotation as identity (kinematics.cpp:110-113, 176-178)
    urdf = """<robot name="r"><link name="base"/><link name="l1"/><link name="l2"/>
    <joint name="f" type="fixed"><parent link="base"/><child link="l1"/><origin xyz="0 0 0" rpy="0.001 0 0"/></joint>
    <joint name="g" type="fixed"><parent link="l1"/><child link="l2"/><origin xyz="0 1 0" rpy="0 0 0"/></joint>
    </robot>"""
    kin = urdf_parse.kinematic_model(urdf)
    p = kin.get_link_pose(kin.get_link_ids(["l2"])[0])
    # the 1 mrad roll is dropped: l1 keeps the base rotation, so l2's offset is
    # not rotated, where exact FK puts l2 at z = sin(0.001)
    assert np.allclose(p[3:], [0, 0, 0, 1])
    assert np.array_equal(p[:3], [0.0, 1.0, 0.0])
    assert abs(numpy_fk(urdf, {})["l2"][2, 3] - math.sin(0.001)) < 1e-15


def test_nan_configuration_is_invalid():
    wl = synth.workload("plain_two_bars", scale=1 / 256)
    cst = oracle_cst(wl)
    q = wl.Q[0].copy()
    q[3] = np.nan
    assert not cst.is_valid(q)


def test_sphere_reorder_is_farthest_point():
    wl = synth.workload("plain_two_bars", scale=1 / 256)
    cst = oracle_cst(wl, reorder=True)
    cst0 = oracle_cst(wl, reorder=False)
    cst.update_kintree(np.zeros(7), True)
    cst0.update_kintree(np.zeros(7), True)
    c1, r1 = cst.get_all_spheres()
    c0, r0 = cst0.get_all_spheres()
    # same multiset of spheres
    assert np.allclose(np.sort(r1), np.sort(r0))
    # within each group: every next sphere is (one of) the farthest from the chosen ones
    start = 0
    for _, pos, radii, _ in wl.robot.sphere_specs:
        k = len(radii)
        g = c1[start:start + k]
        for i in range(1, k):
            chosen, rest = g[:i], g[i:]
            dmin = np.linalg.norm(rest[:, None, :] - chosen[None, :, :], axis=-1).min(1)
            assert dmin[0] >= dmin.max() - 1e-12
        start += k


def _reorder_groups(rng):
    groups = [rng.normal(0.0, 0.1, size=(k, 3)) for k in range(1, 21)]
    groups += [rng.normal(0.0, 0.1, size=(k, 3)) for k in (33, 64, 97)]
    # ties: lattice points (equal distances everywhere) and repeated positions
    grid = np.stack(np.meshgrid(*[np.arange(3.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    groups += [grid, np.zeros((5, 3)), np.repeat(rng.normal(size=(3, 3)), 3, axis=0)]
    return groups


def test_reorder_order_matches_libstdcxx_unordered_set(tmp_path):
    # reorder_spheres_for_informative_collision_check (primitive_sphere_collision.cpp:53-96)
    # run with libstdc++'s real std::unordered_set<size_t>; the oracle models its
    # iteration order with a vector
    import shutil
    import subprocess

    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "reorder_stdset.cpp")
    exe = str(tmp_path / "reorder_stdset")
    subprocess.run([cxx, "-O2", "-std=c++17", src, "-o", exe], check=True)
    groups = _reorder_groups(np.random.default_rng(77))
    text = f"{len(groups)}\n" + "".join(
        f"{len(g)}\n" + "".join(f"{x:.17g} {y:.17g} {z:.17g}\n" for x, y, z in g) for g in groups)
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    for g, line in zip(groups, out):
        std_order = [int(v) for v in line.split()]
        assert sorted(std_order) == list(range(len(g)))
        assert _oracle.informative_order(g) == std_order, f"group of {len(g)}"


def _group_layout(robot):
    """Global sphere index ranges per group (spec order) and, per listed self
    pair, the (i, j) sphere index pairs: every sphere of the first group that
    carries link a against every sphere of the first group that carries link b
    (primitive_sphere_collision.cpp:162-177, all k1 x k2 pairs)."""
    sizes = [len(s[2]) for s in robot.sphere_specs]
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    first = {}
    for gi, s in enumerate(robot.sphere_specs):
        first.setdefault(s[0], gi)
    I, J = [], []
    for a, b in robot.self_pairs:
        ia = np.arange(start[first[a]], start[first[a] + 1])
        ib = np.arange(start[first[b]], start[first[b] + 1])
        I.append(np.repeat(ia, len(ib)))
        J.append(np.tile(ib, len(ia)))
    return start, np.concatenate(I), np.concatenate(J)


def _numpy_centres(robot, q):
    """World sphere centres in spec order from the independent numpy FK."""
    T = numpy_fk(robot.urdf, dict(zip(robot.joint_names, q)))
    return np.concatenate([(T[ln][:3, :3] @ np.asarray(pos, float).T).T + T[ln][:3, 3]
                           for ln, pos, _, _ in robot.sphere_specs])


def _self_brute(robot, q, radii):
    """min over listed sphere pairs of |c1 - c2| - (r1 + r2) (exact FK), and the gap to the runner-up."""
    _, I, J = _group_layout(robot)
    C = _numpy_centres(robot, q)
    d = np.sqrt(((C[I] - C[J]) ** 2).sum(-1)) - (radii[I] + radii[J])
    s = np.sort(d)
    return s[0], s[1] - s[0]


def _dual_arm_configs(n, seed):
    """n dual-arm configurations, a third each with a self collision, with the
    nearest listed sphere pair within the 0.1 m cutoff, and uniform; picked
    from a uniform pool by the oracle's Taylor-FK centres (selection only)."""
    robot = synth.dual_arm(0)
    pool = np.random.default_rng(seed).uniform(-2.9, 2.9, size=(80 * n, 14))
    cst = oracle_cst(synth.Workload("pool", robot, []))
    C = cst.sphere_centers_batch(pool, False)
    _, r = cst.get_all_spheres()
    _, I, J = _group_layout(robot)
    m = (np.sqrt(((C[:, I] - C[:, J]) ** 2).sum(-1)) - (r[I] + r[J])).min(1)
    k = n // 3
    hit, near, far = np.flatnonzero(m < 0), np.flatnonzero((m > 0) & (m < 0.1)), np.flatnonzero(m >= 0.1)
    assert len(hit) >= k and len(near) >= k
    return pool[np.concatenate([hit[:k], near[:k], far[:n - 2 * k]])]


@pytest.mark.parametrize("with_prims", [False, True])
def test_self_collision_is_valid_vs_brute_force(with_prims):
    # check_self_collision (primitive_sphere_collision.cpp:222-254): a listed pair
    # collides iff |c1 - c2|^2 < (r1 + r2)^2; is_valid = ext ok and self ok
    robot = synth.dual_arm(0)
    prims = synth.mixed_primitives(24, seed=5) if with_prims else []
    wl = synth.Workload("dual_arm_self", robot, prims, Q=_dual_arm_configs(300, 11))
    cst = oracle_cst(wl)
    C = cst.sphere_centers_batch(wl.Q, False)  # Taylor FK, as is_valid uses
    _, r = cst.get_all_spheres()
    _, I, J = _group_layout(robot)
    D = C[:, I] - C[:, J]
    sq = D[..., 0] * D[..., 0] + D[..., 1] * D[..., 1] + D[..., 2] * D[..., 2]
    rs = (r[I] + r[J]) ** 2
    self_ok = (sq >= rs).all(1)
    ext_ok = np.ones(len(wl.Q), bool)
    if prims:
        P = [oracle_prim(p) for p in prims]
        for b in range(len(wl.Q)):
            ext = np.array([[p.evaluate(C[b, s]) - r[s] for p in P] for s in range(len(r))])
            ext_ok[b] = bool((ext > 0).all())
    ok = self_ok & ext_ok
    got = cst.is_valid_batch(wl.Q, 4).astype(bool)
    margin = np.abs(sq - rs).min(1)
    keep = margin > 1e-12
    assert keep.mean() > 0.99
    assert np.array_equal(got[keep], ok[keep])
    # both outcomes of the self test occur, including configurations whose only
    # collision is a self collision
    assert 0.2 < self_ok.mean() < 0.8
    if prims:
        assert (~self_ok & ext_ok).any() and (self_ok & ~ext_ok).any()


def test_self_evaluate_value_vs_brute_force():
    # evaluate_self_collision (primitive_sphere_collision.cpp:324-383): min over
    # listed sphere pairs of |c1 - c2| - (r1 + r2), capped at the 0.1 cutoff
    robot = synth.dual_arm(0)
    wl = synth.Workload("dual_arm_self", robot, [], Q=_dual_arm_configs(120, 12))
    cst = oracle_cst(wl, reorder=False)
    assert cst.cst_dim() == 1
    radii = np.concatenate([np.asarray(s[2], float) for s in robot.sphere_specs])
    vals, _ = cst.evaluate_batch(wl.Q, 4)
    below = 0
    for b, q in enumerate(wl.Q):
        dmin, _ = _self_brute(robot, q, radii)
        assert abs(vals[b, 0] - min(dmin, 0.1)) < 1e-12
        below += dmin < 0.1
    assert 0.5 < below / len(wl.Q) < 0.9
    # reordering the spheres does not change the minimum
    v2, _ = oracle_cst(wl, reorder=True).evaluate_batch(wl.Q, 4)
    assert np.allclose(v2, vals, atol=1e-14)


def test_self_evaluate_row_vs_central_fd():
    # the self row (c1 - c2)^T / |c1 - c2| (J1 - J2) (:375-381) against FP64
    # central differences of the brute-force minimum, away from argmin switches
    robot = synth.dual_arm(0)
    wl = synth.Workload("dual_arm_self", robot, [], Q=_dual_arm_configs(200, 13))
    cst = oracle_cst(wl)
    radii = np.concatenate([np.asarray(s[2], float) for s in robot.sphere_specs])
    h = 1e-6
    checked = 0
    for q in wl.Q:
        dmin, gap = _self_brute(robot, q, radii)
        if not (dmin < 0.09 and gap > 1e-4):
            continue
        v, Jrow = cst.evaluate(q)
        fd = np.array([(_self_brute(robot, q + h * e, radii)[0] - _self_brute(robot, q - h * e, radii)[0]) / (2 * h)
                       for e in np.eye(14)])
        np.testing.assert_allclose(Jrow[0], fd, rtol=1e-6, atol=1e-8)
        assert np.any(np.abs(fd) > 1e-3)
        checked += 1
        if checked == 12:
            break
    assert checked >= 8


def test_self_two_spheres_closed_form():
    # one revolute link about z carries a sphere at (1, 0, 0); the base carries
    # one at (bx, by, 0): d(q) = |(cos q - bx, sin q - by)| - (ra + rb),
    # dd/dq = ((cos q - bx)(-sin q) + (sin q - by) cos q) / |.|
    urdf = _one_link_robot()
    ra, rb, bx, by = 0.2, 0.15, 0.9, 0.6
    spec_a = _oracle.SphereSpec("l1", np.array([[1.0, 0.0, 0.0]]), [ra], False)
    spec_b = _oracle.SphereSpec("base", np.array([[bx, by, 0.0]]), [rb], False)
    cst = _oracle.SphereCollisionCst(urdf_parse.kinematic_model(urdf), ["j1"], _oracle.BaseType.FIXED,
                                     [spec_a, spec_b], [("l1", "base")], [], False)
    for q in np.linspace(-1.0, 2.0, 31):
        dx, dy = math.cos(q) - bx, math.sin(q) - by
        n = math.hypot(dx, dy)
        d = n - (ra + rb)
        g = (dx * -math.sin(q) + dy * math.cos(q)) / n
        v, Jrow = cst.evaluate(np.array([q]))
        if d < 0.1:
            assert abs(v[0] - d) < 1e-14 and abs(Jrow[0, 0] - g) < 1e-12
        else:
            assert v[0] == 0.1 and Jrow[0, 0] == 0.0
        if abs(n * n - (ra + rb) ** 2) > 1e-4:  # Taylor FK moves the centre by < 3e-5
            assert cst.is_valid(np.array([q])) == (n * n >= (ra + rb) ** 2)


# ------------------------------------------------------ motion validators
def test_sequence_table_known_answers():
    assert _oracle.farthest_sequence(8) == [0, 7, 4, 2, 6, 5, 3, 1]  # motion_validator.cpp:23-37
    assert _oracle.farthest_sequence(20) == [0, 19, 10, 5, 15, 17, 13, 8, 3, 18, 16, 14, 12, 11, 9, 7, 6, 4, 2, 1]
    table = _oracle.sequence_table()
    for n in range(1, 129):  # generate_sequence.py:66-73
        seq = table[n - 1][:n]
        assert sorted(seq) == list(range(n)) and seq[0] == 0
        if n > 1:
            assert seq[1] == n - 1
    rates = _oracle.motion_travel_rates(0.15)
    assert np.allclose(rates[1:], [0.6, 0.3, 0.9, 0.75, 0.45, 0.15]) and rates[0] < 0


def test_step_ratio_closed_forms():
    w = [0.05] * 7
    s1 = [0.0] * 7
    s2 = [0.0, 0.3, -0.9, 0.0, 0.1, 0.0, 0.0]
    assert _oracle.box_step_ratio(s1, s2, w) == pytest.approx(0.05 / 0.9)
    assert _oracle.box_step_ratio(s1, [5e-7] * 7, w) == math.inf
    assert _oracle.euclid_step_ratio(s1, s2, 0.1) == pytest.approx(0.1 / np.linalg.norm(s2))
    n_linear = len(_oracle.motion_travel_rates(0.005))  # n_test = 202 > 128: linear branch
    assert n_linear == 201


def test_check_motion_matches_waypoint_enumeration():
    wl = synth.workload("edges_mixed16", scale=1 / 128)
    cst = oracle_cst(wl)
    ok, n = cst.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 4)
    for e in range(len(ok)):
        r = _oracle.box_step_ratio(list(wl.S1[e]), list(wl.S2[e]), list(wl.box_width))
        states = [wl.S1[e] + t * (wl.S2[e] - wl.S1[e]) if t >= 0 else wl.S2[e]
                  for t in _oracle.motion_travel_rates(r)]
        assert bool(ok[e]) == all(cst.is_valid(s) for s in states)


def test_smoothness_cost_matches_matrix_form():
    # smoothcost_fullmat (sqp_based_solver.py:74-96): x^T kron(A, W^2) x
    T, dof = 9, 3
    w = [1.0, 0.5, 2.0]
    acc = np.array([[1, -2, 1], [-2, 4, -2], [1, -2, 1]])
    A = np.zeros((T, T))
    for i in range(1, T - 1):
        A[i - 1:i + 2, i - 1:i + 2] += acc
    P = np.kron(A, np.diag(w) ** 2)
    x = RNG.normal(size=T * dof)
    cost, grad = _oracle.smooth_cost(x, T, dof, w)
    assert cost == pytest.approx(x @ P @ x)
    assert np.allclose(grad, 2 * P @ x)


def _csc_dense(colptr, rowind, values, shape):
    M = np.zeros(shape)
    for j in range(shape[1]):
        for k in range(colptr[j], colptr[j + 1]):
            M[rowind[k], j] = values[k]
    return M


def test_seq_ineq_known_answers_and_structure():
    # SequentialCst with the collision constraint at every waypoint and the
    # motion-step box (tests/test_constraints.py:273-318): row count
    # cst_dim * T + 2 dof (T - 1); msbox rows all >= 0 for steps of half the
    # box, and exactly half of them negative for steps of 1.1 boxes; the
    # sparsity pattern does not change between evaluations.
    wl = synth.workload("grad_mixed32", scale=1 / 256)
    oc = oracle_cst(wl)
    D, C, T = oc.q_dim(), oc.cst_dim(), 4
    msbox = list(np.linspace(0.1, 0.8, D))
    q1 = wl.Q[0]
    for step, check in ((0.5, "ok"), (1.1, "ng")):
        X = np.array([q1 + k * step * np.array(msbox) for k in range(T)])
        c, cp, ri, va, (rows, cols) = _oracle.seq_ineq_evaluate(oc, X, msbox)
        assert rows == C * T + 2 * D * (T - 1) and cols == D * T and len(c) == rows
        tail = c[-2 * D * (T - 1):]
        if check == "ok":
            assert np.all(tail >= 0)
        else:
            assert np.sum(tail < 0) == D * (T - 1) and np.sum(tail > 0) == D * (T - 1)
    X1 = wl.Q[:T]
    X2 = wl.Q[T:2 * T]
    a = _oracle.seq_ineq_evaluate(oc, X1, msbox)
    b = _oracle.seq_ineq_evaluate(oc, X2, msbox)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    # without the box: only the collision rows, C nonzeros per column
    c, cp, ri, va, (rows, cols) = _oracle.seq_ineq_evaluate(oc, X1, None)
    assert rows == C * T and np.array_equal(np.diff(cp), np.full(D * T, C))


def test_seq_ineq_jacobian_vs_central_fd():
    # the CSC Jacobian against FP64 central differences of the stacked values
    # (collision rows use the reference's forward-difference SDF gradient, so
    # compare where the argmin pair is stable)
    wl = synth.workload("grad_mixed32", scale=1 / 256)
    oc = oracle_cst(wl)
    D, T = oc.q_dim(), 3
    msbox = [0.2] * D
    X = wl.Q[:T]
    c, cp, ri, va, shape = _oracle.seq_ineq_evaluate(oc, X, msbox)
    J = _csc_dense(cp, ri, va, shape)
    h = 1e-6
    x = X.ravel()
    fd = np.empty(shape)
    fd4 = np.empty(shape)
    for j in range(len(x)):
        e = np.zeros(len(x))
        e[j] = 1.0
        cp_ = _oracle.seq_ineq_evaluate(oc, (x + h * e).reshape(T, D), msbox)[0]
        cm_ = _oracle.seq_ineq_evaluate(oc, (x - h * e).reshape(T, D), msbox)[0]
        cp4 = _oracle.seq_ineq_evaluate(oc, (x + 4 * h * e).reshape(T, D), msbox)[0]
        cm4 = _oracle.seq_ineq_evaluate(oc, (x - 4 * h * e).reshape(T, D), msbox)[0]
        fd[:, j] = (cp_ - cm_) / (2 * h)
        fd4[:, j] = (cp4 - cm4) / (8 * h)
    smooth = np.abs(fd - fd4) <= 1e-6
    assert smooth.mean() > 0.9
    assert np.all(np.abs(J - fd)[smooth] <= 1e-4 * np.maximum(np.abs(fd[smooth]), 1.0))
