"""Batched multi-start inverse kinematics on the device path (src/plainmp/ik.py:42-156).

The reference's ``solve_ik`` runs SLSQP from one random seed at a time, up to
``max_trial`` times, and accepts the first result whose squared pose error
is at most ``acceptable_error`` and which the inequality constraint reports
valid; its expensive inner calls are the pose and collision Jacobians.  Here
every restart runs at once: each iteration makes one link-pose launch (FP64
FK), one link-Jacobian launch (four with a rotation target), one evaluate
launch (collision values and rows) and one step launch for all restarts
(pmp_ik_step_batch: a damped Gauss-Newton step per restart on the pose
residual with the collision values below ``margin`` as penalty residuals,
its normal equations solved in registers, projected onto the joint bounds;
robots with more than 16 q columns take the same step in torch).  The
answer is the first restart, in restart order, that passes the reference's
acceptance test (squared pose residual <= ``acceptable_error``,
``is_valid``), so a result is reproducible from the seed.

Targets are ``LinkPoseCst`` poses (constraints/primitive_link_pose.cpp:32-57):
a 3-vector position, a 6-vector position + roll-pitch-yaw (the residual is
``getRPY() - rpy``, transform.hpp:93-120) or a 7-vector position + quaternion
coefficients x, y, z, w (residual ``quat.coeffs() - q``).  The rotation rows
of the Jacobian are the reference's (kinematics/algorithm.cpp:21-47, 255-272):
``rpy_derivative(rpy, w_j)`` and ``q_derivative(q^-1, w_j)`` of each revolute
joint's world axis w_j, which this module obtains from link-Jacobian launches
at the link origin and at its three unit axes: the point Jacobian column of a
revolute joint is w_j x (p - o_j), so the differences D_k at p + u_k give
w_j = 1/2 sum_k u_k x D_k for the orthonormal world axes u_k of the link
(prismatic and unrelated joints give zero).
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np


@dataclass
class IKConfig:
    acceptable_error: float = 1e-6  # squared pose residual, as solve_ik's res.fun test
    n_iter: int = 60
    damping: float = 1e-4
    max_step: float = 0.3  # largest joint step per iteration (rad or m)
    margin: float = 0.005  # collision values below this are penalised
    penalty: float = 10.0


@dataclass
class IKResult:
    q: np.ndarray
    elapsed_time: float
    success: bool
    n_trial: int


def _rpy_of(qx, qy, qz, qw):
    """QuatTrans::getRPY (kinematics/transform.hpp:93-120), elementwise."""
    import torch

    sqx, sqy, sqz, sqw = qx * qx, qy * qy, qz * qz, qw * qw
    sarg = -2 * (qx * qz - qw * qy)
    half_pi = 1.57079632679489661923
    lock = 2 * torch.atan2(qx, qy)
    pitch = torch.where(sarg <= -0.99999, torch.full_like(sarg, -half_pi),
                        torch.where(sarg >= 0.99999, torch.full_like(sarg, half_pi),
                                    torch.asin(torch.clamp(sarg, -1.0, 1.0))))
    roll = torch.where((sarg <= -0.99999) | (sarg >= 0.99999), torch.zeros_like(sarg),
                       torch.atan2(2 * (qy * qz + qw * qx), sqw - sqx - sqy + sqz))
    yaw = torch.where(sarg <= -0.99999, -lock, torch.where(sarg >= 0.99999, lock,
                                                            torch.atan2(2 * (qx * qy + qw * qz), sqw + sqx - sqy - sqz)))
    return torch.stack([roll, pitch, yaw], dim=-1)


def _rotation_rows(kind, quat, rpy, w):
    """Rotation rows of the Jacobian, (R, 3 or 4, D), from the world axes
    w (R, 3, D): rpy_derivative or q_derivative of the inverse quaternion
    (kinematics/algorithm.cpp:21-47)."""
    import torch

    wx, wy, wz = w[:, 0], w[:, 1], w[:, 2]
    if kind == 6:
        a2, a3 = -rpy[:, 1:2], -rpy[:, 2:3]
        c2, c3, s2, s3 = torch.cos(a2), torch.cos(a3), torch.sin(a2), torch.sin(a3)
        dr = c3 / c2 * wx - s3 / c2 * wy
        dp = s3 * wx + c3 * wy
        dy = -c3 * s2 / c2 * wx + s3 * s2 / c2 * wy + wz
        return torch.stack([dr, dp, dy], dim=1)
    # q^-1 of a unit quaternion: (-x, -y, -z, w)
    x, y, z, qw = (-quat[:, 0:1], -quat[:, 1:2], -quat[:, 2:3], quat[:, 3:4])
    dxdt = 0.5 * (wz * y - wy * z + wx * qw)
    dydt = 0.5 * (-wz * x + wx * z + wy * qw)
    dzdt = 0.5 * (wy * x - wx * y + wz * qw)
    dwdt = 0.5 * (-wx * x - wy * y - wz * z)
    return torch.stack([dxdt, dydt, dzdt, -dwdt], dim=1)  # coeffs of Quaternion(-dwdt, dxdt, dydt, dzdt)


def _rot_matrix(quat):
    """World rotation matrices (R, 3, 3) of unit quaternions (x, y, z, w)."""
    import torch

    x, y, z, w = quat[:, 0], quat[:, 1], quat[:, 2], quat[:, 3]
    return torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], dim=-1),
        torch.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], dim=-1),
        torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], dim=-1)], dim=1)


def solve_ik_batched(cst, link_name: str, target: Sequence[float], lb: Sequence[float], ub: Sequence[float], *,
                     check_collision: bool = True, q_seed: Optional[np.ndarray] = None, n_restarts: int = 256,
                     seed: int = 0, config: Optional[IKConfig] = None) -> IKResult:
    """IK for ``link_name`` of the constraint's robot to a 3-, 6- or 7-vector
    ``target`` (position, + RPY, + quaternion x, y, z, w); ``cst`` also
    supplies the collision constraint when ``check_collision`` is set.
    ``q_seed`` (optional) is the first restart; the others are uniform in
    [lb, ub] from ``seed``."""
    import torch

    ts = time.time()
    kind = len(target)
    if kind not in (3, 6, 7):
        raise RuntimeError("All poses must be 3 or 6 or 7 dimensional")
    cfg = config or IKConfig()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    kin = cst.get_kin()
    link = int(kin.get_link_ids([link_name])[0])
    n_nodes = int(kin.n_links)
    D, C, R = cst.q_dim(), cst.cst_dim(), int(n_restarts)
    lo = torch.tensor(np.asarray(lb, float), device=dev, dtype=torch.float64)
    hi = torch.tensor(np.asarray(ub, float), device=dev, dtype=torch.float64)
    tgt = torch.tensor(np.asarray(target, float), device=dev, dtype=torch.float64)
    gen = torch.Generator(device=dev).manual_seed(seed)
    Q = lo + (hi - lo) * torch.rand((R, D), device=dev, dtype=torch.float64, generator=gen)
    if q_seed is not None:
        Q[0] = torch.tensor(np.asarray(q_seed, float), device=dev, dtype=torch.float64)
    poses = torch.empty((n_nodes * 7, R), device=dev, dtype=torch.float64)
    J = torch.empty((3, D, R), device=dev, dtype=torch.float32)
    Ju = [torch.empty((3, D, R), device=dev, dtype=torch.float32) for _ in range(3 if kind > 3 else 0)]
    units = ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])
    vals = torch.empty((max(C, 1), R), device=dev, dtype=torch.float32)
    cjac = torch.empty((max(C, 1), D, R), device=dev, dtype=torch.float32)
    eye = torch.eye(D, device=dev, dtype=torch.float64)
    use_col = check_collision and C > 0

    def pose_error(qs):
        """Residual (R, kind) of LinkPoseCst, the link quaternion and RPY."""
        cst._dev_link_poses(qs.data_ptr(), R, R, 1, poses.data_ptr(), sh)
        pose = poses[7 * link:7 * link + 7].T  # (R, 7): x, y, z, qx, qy, qz, qw
        quat = pose[:, 3:7]
        rpy = _rpy_of(quat[:, 0], quat[:, 1], quat[:, 2], quat[:, 3]) if kind == 6 else None
        e = pose[:, :3] - tgt[:3]
        if kind == 6:
            e = torch.cat([e, rpy - tgt[3:]], dim=1)
        elif kind == 7:
            e = torch.cat([e, quat - tgt[3:]], dim=1)
        return e, quat, rpy

    if D <= 16 and cfg.n_iter > 0:
        # fused: per iteration the Jacobian launches (the first also giving the
        # link pose), the evaluate launch and one step kernel (the normal
        # equations solved per restart in registers), one iteration captured
        # into a HIP graph and replayed n_iter times
        from ._native.ompl import _dev_capture_ik_graph

        qs = Q.T.contiguous()
        C_used = C if use_col else 0
        graph = _dev_capture_ik_graph(cst, qs.data_ptr(), R, kind, link, tgt.data_ptr(),
                                      J.data_ptr(), [u.data_ptr() for u in Ju], vals.data_ptr(), cjac.data_ptr(),
                                      C_used, lo.data_ptr(), hi.data_ptr(), cfg.damping, cfg.max_step, cfg.margin,
                                      float(cfg.penalty), 1, sh)
        # the graph replays on the context's stream, or on a stream of its own
        # when the caller's is the null stream: the inputs are made ready
        # first, and destroy() waits for the replays
        stream.synchronize()
        for _ in range(cfg.n_iter):
            graph.launch()
        graph.destroy()
        Q = qs.T.contiguous()
    for _ in range(cfg.n_iter if D > 16 else 0):
        qs = Q.T.contiguous()
        e, quat, rpy = pose_error(qs)
        cst._dev_link_jacobian(qs.data_ptr(), R, R, link, [], J.data_ptr(), sh)
        A = J.permute(2, 0, 1).double()  # (R, 3, D)
        if kind > 3:
            U = _rot_matrix(quat)  # columns: the link's world axes
            w = torch.zeros((R, 3, D), device=dev, dtype=torch.float64)
            for k in range(3):
                cst._dev_link_jacobian(qs.data_ptr(), R, R, link, units[k], Ju[k].data_ptr(), sh)
                Dk = Ju[k].permute(2, 0, 1).double() - A  # (R, 3, D) = w_j x u_k
                uk = U[:, :, k:k + 1].expand(-1, -1, D)
                w += 0.5 * torch.linalg.cross(uk, Dk, dim=1)
            A = torch.cat([A, _rotation_rows(kind, quat, rpy, w)], dim=1)
        r = e
        if use_col:
            cst._dev_evaluate(qs.data_ptr(), R, R, vals.data_ptr(), cjac.data_ptr(), sh)
            v = vals[:C].T.double()  # (R, C)
            act = (v < cfg.margin).double()
            w = np.sqrt(cfg.penalty)
            A = torch.cat([A, -w * cjac[:C].permute(2, 0, 1).double() * act[..., None]], dim=1)
            r = torch.cat([e, w * (cfg.margin - v) * act], dim=1)
        At = A.transpose(1, 2)
        # A^T A + damping I is positive definite: solve_ex skips the error
        # check, whose read of the result on the host would stall every iteration
        step = -torch.linalg.solve_ex(At @ A + cfg.damping * eye, (At @ r[..., None]))[0][..., 0]
        norm = step.abs().max(dim=1, keepdim=True).values
        step = step * torch.clamp(cfg.max_step / torch.clamp(norm, min=1e-300), max=1.0)
        Q = torch.minimum(torch.maximum(Q + step, lo), hi)

    qs = Q.T.contiguous()
    err2 = (pose_error(qs)[0] ** 2).sum(dim=1)
    ok = err2 <= cfg.acceptable_error
    if check_collision:
        valid = torch.empty(R, device=dev, dtype=torch.uint8)
        cst._dev_is_valid(qs.data_ptr(), R, R, valid.data_ptr(), 0, sh)
        ok = ok & (valid == 1)
    hits = torch.nonzero(ok).flatten()
    if len(hits) == 0:
        return IKResult(np.empty([0]), time.time() - ts, False, R)
    i = int(hits[0].item())
    return IKResult(Q[i].cpu().numpy(), time.time() - ts, True, i + 1)


__all__ = ["IKConfig", "IKResult", "solve_ik_batched"]
