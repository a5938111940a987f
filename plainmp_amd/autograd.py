"""torch.autograd bindings of the collision distances.

``pair_clearance(cst, q)``: signed distance of every (sphere, primitive)
pair, shape (B, S, P), differentiable in q (B, q_dim).  The backward pass
is the hand-written adjoint kernel (``pmp_pair_clearance_vjp``): per sphere
the upstream weights fold into one world-space force, which each joint
column receives through a x (c - o).  No (B, S, P, q_dim) tensor is formed.

``constraint_values(cst, q)``: the reference-semantics values of
``SphereCollisionCst::evaluate`` (min external clearance, min self clearance,
0.1 m cutoff), shape (B, cst_dim); its backward uses the Jacobian rows the
evaluate kernel returns.  This is the Jacobian that IK and SQP consume
(constraints/primitive_sphere_collision.cpp:256-399).

Both run on the caller's current CUDA (HIP) stream; q may be float32 or
float64 (configurations enter the kernels in float64).
"""

from __future__ import annotations

import torch


def _prep(q: torch.Tensor, cst):
    if not q.is_cuda:
        raise ValueError("q must be a GPU tensor")
    if q.dim() != 2 or q.shape[1] != cst.q_dim():
        raise ValueError(f"q must be (B, {cst.q_dim()})")
    return q.detach().to(torch.float64).t().contiguous()  # (D, B) structure of arrays


class PairClearanceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, cst):
        qd = _prep(q, cst)
        B = q.shape[0]
        S, P = cst.n_spheres(), cst.n_prims()
        dist = torch.empty((S, P, B), dtype=torch.float32, device=q.device)
        if B and S and P:
            cst._dev_pair_clearance(qd.data_ptr(), B, B, dist.data_ptr(), 0, torch.cuda.current_stream().cuda_stream)
        ctx.save_for_backward(qd)
        ctx.cst = cst
        ctx.q_dtype = q.dtype
        return dist.permute(2, 0, 1)

    @staticmethod
    def backward(ctx, grad):
        (qd,) = ctx.saved_tensors
        cst = ctx.cst
        D, B = qd.shape
        gd = grad.permute(1, 2, 0).to(torch.float32).contiguous()  # (S, P, B)
        gq = torch.empty((D, B), dtype=torch.float32, device=qd.device)
        if B:
            cst._dev_pair_clearance_vjp(qd.data_ptr(), B, B, gd.data_ptr(), gq.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream)
        return gq.t().to(ctx.q_dtype), None


class ConstraintValuesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, cst):
        qd = _prep(q, cst)
        B, D, C = q.shape[0], cst.q_dim(), cst.cst_dim()
        vals = torch.empty((C, B), dtype=torch.float32, device=q.device)
        jac = torch.empty((C, D, B), dtype=torch.float32, device=q.device)
        if B and C:
            cst._dev_evaluate(qd.data_ptr(), B, B, vals.data_ptr(), jac.data_ptr(),
                              torch.cuda.current_stream().cuda_stream)
        ctx.save_for_backward(jac)
        ctx.q_dtype = q.dtype
        return vals.t()

    @staticmethod
    def backward(ctx, grad):
        (jac,) = ctx.saved_tensors  # (C, D, B)
        g = (grad.to(torch.float32).t()[:, None, :] * jac).sum(0)  # (D, B), elementwise only
        return g.t().to(ctx.q_dtype), None


def pair_clearance(cst, q: torch.Tensor) -> torch.Tensor:
    """(B, S, P) signed distances sdf_p(c_s) - r_s with exact FK."""
    return PairClearanceFn.apply(q, cst)


def constraint_values(cst, q: torch.Tensor) -> torch.Tensor:
    """(B, cst_dim) reference-semantics collision values (ext first, then self)."""
    return ConstraintValuesFn.apply(q, cst)


def jacobian_rows(cst, q: torch.Tensor):
    """(values (B, cst_dim), Jacobian (B, cst_dim, q_dim)) without autograd."""
    qd = _prep(q, cst)
    B, D, C = q.shape[0], cst.q_dim(), cst.cst_dim()
    vals = torch.empty((C, B), dtype=torch.float32, device=q.device)
    jac = torch.empty((C, D, B), dtype=torch.float32, device=q.device)
    if B and C:
        cst._dev_evaluate(qd.data_ptr(), B, B, vals.data_ptr(), jac.data_ptr(), torch.cuda.current_stream().cuda_stream)
    return vals.t(), jac.permute(2, 0, 1)


__all__ = ["pair_clearance", "constraint_values", "jacobian_rows", "PairClearanceFn", "ConstraintValuesFn"]
