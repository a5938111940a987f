"""Batched multi-start inverse kinematics on the device path (src/plainmp/ik.py:42-156).

The reference's ``solve_ik`` runs SLSQP from one random seed at a time, up to
``max_trial`` times, and accepts the first result whose squared pose error
is at most ``acceptable_error`` and which the inequality constraint reports
valid; its expensive inner calls are the pose and collision Jacobians.  Here
every restart runs at once: each iteration makes one link-pose launch (FP64
FK), one link-Jacobian launch and one evaluate launch (collision values and
rows) for all restarts, and takes a damped Gauss-Newton step per restart on
the position error with the collision values below ``margin`` as penalty
residuals, projected onto the joint bounds.  The answer is the first restart,
in restart order, that passes the reference's acceptance test (position
error squared <= ``acceptable_error``, ``is_valid``), so a result is
reproducible from the seed.  Targets are link positions (``LinkPoseCst``
with 3-vector poses, constraints/primitive_link_pose.cpp:32-44).
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np


@dataclass
class IKConfig:
    acceptable_error: float = 1e-6  # squared position error, as solve_ik's res.fun test
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


def solve_ik_batched(cst, link_name: str, target: Sequence[float], lb: Sequence[float], ub: Sequence[float], *,
                     check_collision: bool = True, q_seed: Optional[np.ndarray] = None, n_restarts: int = 256,
                     seed: int = 0, config: Optional[IKConfig] = None) -> IKResult:
    """Position IK for ``link_name`` of the constraint's robot; ``cst`` also
    supplies the collision constraint when ``check_collision`` is set.
    ``q_seed`` (optional) is the first restart; the others are uniform in
    [lb, ub] from ``seed``."""
    import torch

    ts = time.time()
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
    vals = torch.empty((max(C, 1), R), device=dev, dtype=torch.float32)
    cjac = torch.empty((max(C, 1), D, R), device=dev, dtype=torch.float32)
    eye = torch.eye(D, device=dev, dtype=torch.float64)
    use_col = check_collision and C > 0

    def position_error(qs):
        cst._dev_link_poses(qs.data_ptr(), R, R, 1, poses.data_ptr(), sh)
        return poses[7 * link:7 * link + 3].T - tgt  # (R, 3)

    for _ in range(cfg.n_iter):
        qs = Q.T.contiguous()
        e = position_error(qs)
        cst._dev_link_jacobian(qs.data_ptr(), R, R, link, [], J.data_ptr(), sh)
        A = J.permute(2, 0, 1).double()  # (R, 3, D)
        r = e
        if use_col:
            cst._dev_evaluate(qs.data_ptr(), R, R, vals.data_ptr(), cjac.data_ptr(), sh)
            v = vals[:C].T.double()  # (R, C)
            act = (v < cfg.margin).double()
            w = np.sqrt(cfg.penalty)
            A = torch.cat([A, -w * cjac[:C].permute(2, 0, 1).double() * act[..., None]], dim=1)
            r = torch.cat([e, w * (cfg.margin - v) * act], dim=1)
        At = A.transpose(1, 2)
        step = -torch.linalg.solve(At @ A + cfg.damping * eye, (At @ r[..., None]))[..., 0]
        norm = step.abs().max(dim=1, keepdim=True).values
        step = step * torch.clamp(cfg.max_step / torch.clamp(norm, min=1e-300), max=1.0)
        Q = torch.minimum(torch.maximum(Q + step, lo), hi)

    qs = Q.T.contiguous()
    err2 = (position_error(qs) ** 2).sum(dim=1)
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
