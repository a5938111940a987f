"""Sphere collision constraint (plainmp.constraint; constraints/primitive_sphere_collision.hpp).

``SphereCollisionCst`` keeps the reference's single-configuration methods
(``is_valid``, ``evaluate``, ``update_kintree``, ``get_all_spheres``,
``get_group_spheres``) and adds batched ones (``is_valid_batch``,
``evaluate_batch``, ``sphere_clearance_batch``, ``sphere_centers_batch``,
``pair_clearance_batch``, ``traj_cost_grad_batch``, ``traj_sqp_ineq``).
"""

from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._native.constraint import SphereAttachmentSpec, SphereCollisionCst  # noqa: F401
from ._native.ompl import _dev_capture_traj_graph


def capture_traj_graph(cst: SphereCollisionCst, X_ptr: int, ldx: int, N: int, offsets: Sequence[int],
                       weights: Sequence[float], vals_ptr: int, jac_ptr: int, cost_ptr: int, grad_ptr: int,
                       stream: int):
    """``traj_cost_grad`` over device buffers (pointers as integers, layouts of
    pmp_traj_cost_grad_batch) captured once into a HIP graph: ``launch()``
    replays every trajectory's evaluate rows, smoothness cost and gradient in
    one launch; ``destroy()`` frees it (the constraint outlives the graph)."""
    return _dev_capture_traj_graph(cst, X_ptr, ldx, N, [int(o) for o in offsets], [float(w) for w in weights],
                                   vals_ptr, jac_ptr, cost_ptr, grad_ptr, stream)


def sqp_ineq_matrices(cst: SphereCollisionCst, X: np.ndarray, offsets: Sequence[int],
                      msbox: Optional[Sequence[float]] = None) -> List[Tuple[np.ndarray, "object"]]:
    """The SQP inequality sequence of every trajectory, as the reference's
    ``SequentialCst.evaluate`` returns it (constraints/sequential_constraint.cpp:63-127):
    a list of (values, scipy.sparse.csc_matrix Jacobian), one per trajectory of
    ``X`` ((N, q_dim) waypoints, trajectory i = ``X[offsets[i]:offsets[i+1]]``),
    with the collision constraint at every waypoint and, when ``msbox`` is
    given, the motion-step box between consecutive waypoints.  All
    trajectories are assembled on the GPU in one launch."""
    import scipy.sparse as sp

    D = cst.q_dim()
    c, colptr, rowind, values, row_off, nnz_off, cp_off = cst.traj_sqp_ineq(
        np.ascontiguousarray(X, dtype=np.float64), [int(o) for o in offsets],
        [] if msbox is None else [float(w) for w in msbox])
    out = []
    for i in range(len(offsets) - 1):
        T = int(offsets[i + 1] - offsets[i])
        rows = int(row_off[i + 1] - row_off[i])
        z0, z1 = int(nnz_off[i]), int(nnz_off[i + 1])
        cp = colptr[cp_off[i]:cp_off[i + 1]]
        J = sp.csc_matrix((values[z0:z1], rowind[z0:z1], cp), shape=(rows, D * T))
        out.append((c[row_off[i]:row_off[i + 1]], J))
    return out


__all__ = ["SphereAttachmentSpec", "SphereCollisionCst", "sqp_ineq_matrices", "capture_traj_graph"]
