"""torch.autograd path: the adjoint (VJP) kernel against the full per-pair
Jacobian tensor and FP64 central differences, and the reference-semantics
values' backward against the evaluate kernel's rows."""

import numpy as np
import pytest

from plainmp_amd import synth
from support import oracle_cst, product_cst

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["grad_mixed32", "cloud_shelves"])
def test_pair_clearance_vjp_matches_jacobian_tensor(name):
    import torch

    from plainmp_amd.autograd import pair_clearance

    wl = synth.workload(name, scale=128 / (16384 if name == "cloud_shelves" else 8192))  # 128 configs
    pc = product_cst(wl)
    q = torch.tensor(wl.Q, device="cuda", dtype=torch.float64, requires_grad=True)
    d = pair_clearance(pc, q)
    dist, grad = pc.pair_clearance_batch(wl.Q, True)
    assert np.allclose(d.detach().cpu().numpy(), dist, atol=0, rtol=0)
    w = torch.randn(d.shape, device="cuda", dtype=torch.float32, generator=torch.Generator("cuda").manual_seed(0))
    (d * w).sum().backward()
    ref = np.einsum("bsp,bspj->bj", w.cpu().numpy().astype(np.float64), grad.astype(np.float64))
    got = q.grad.cpu().numpy()
    scale = np.abs(w.cpu().numpy()).sum(axis=(1, 2))[:, None] * np.abs(grad).max()
    print("vjp max rel", (np.abs(got - ref) / scale).max())
    assert (np.abs(got - ref) <= 1e-5 * scale).all()


def test_pair_clearance_vjp_vs_fp64_central_fd():
    import torch

    from plainmp_amd.autograd import pair_clearance

    wl = synth.workload("grad_mixed32", scale=1 / 256)  # 32 configs
    oc, pc = oracle_cst(wl), product_cst(wl)
    q = torch.tensor(wl.Q, device="cuda", dtype=torch.float64, requires_grad=True)
    rng = np.random.default_rng(3)
    w = rng.normal(size=(len(wl.Q), pc.n_spheres(), pc.n_prims()))
    (pair_clearance(pc, q) * torch.tensor(w, device="cuda", dtype=torch.float32)).sum().backward()
    got = q.grad.cpu().numpy()

    def loss(Q):
        return (oc.pair_clearance_batch(Q, True, 8) * w).sum(axis=(1, 2))

    D = wl.Q.shape[1]
    fd, fd4 = np.empty_like(got), np.empty_like(got)
    for j in range(D):
        e = np.zeros(D)
        e[j] = 1.0
        fd[:, j] = (loss(wl.Q + 1e-6 * e) - loss(wl.Q - 1e-6 * e)) / 2e-6
        fd4[:, j] = (loss(wl.Q + 4e-6 * e) - loss(wl.Q - 4e-6 * e)) / 8e-6
    smooth = np.abs(fd - fd4) <= 1e-5 * np.maximum(np.abs(fd), 1.0)
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 1.0)
    print("vjp vs fd: smooth", smooth.sum(), "of", fd.size, "max rel", rel[smooth].max())
    assert rel[smooth].max() <= 1e-4


def test_constraint_values_backward_uses_evaluate_rows():
    import torch

    from plainmp_amd.autograd import constraint_values

    wl = synth.workload("grad_mixed32", scale=1 / 64)
    pc = product_cst(wl)
    q = torch.tensor(wl.Q, device="cuda", dtype=torch.float64, requires_grad=True)
    v = constraint_values(pc, q)
    vals, jac = pc.evaluate_batch(wl.Q)
    assert np.allclose(v.detach().cpu().numpy(), vals, atol=0)
    v.sum().backward()
    assert np.allclose(q.grad.cpu().numpy(), jac.sum(axis=1), atol=1e-6)
