"""Parity at the benchmark sizes (GPU): the HIP path against the FP64 oracle
(on every host thread) on the full seeded batches of the five workloads.

* edges_mixed16: 4,096 edges under the autotuned schedule and as replays of
  the captured graph; validity and the reference-order checked counts exact.
* cloud_shelves: 16,384 configurations against the 65,536-point cloud;
  validity exact; batches above the unit pass's 65,536 (plain_two_bars
  81,920, cloud_shelves 69,632) through the lane-per-configuration kernel.
* grad_mixed32: 8,192 configurations; evaluate values within 5e-6 m and rows
  within 1e-4 (relative, unambiguous argmin), every pair distance within 5e-6.
* traj_dual_arm: 2,048 trajectories (122,138 waypoints); evaluate values and
  rows as above, smoothness cost and gradient of every trajectory.
"""

import os

import numpy as np
import pytest

from oracle import _oracle
from plainmp_amd import synth
from support import oracle_cst, product_cst, rows_agree

pytestmark = pytest.mark.gpu

THREADS = max(1, len(os.sched_getaffinity(0)))


def _report(name, **kv):
    print(f"[{name}] " + " ".join(f"{k}={v}" for k, v in kv.items()), flush=True)


def _second_gap(d):
    s = np.sort(d.reshape(len(d), -1), axis=1)
    return s[:, 1] - s[:, 0]


def test_edges_mixed16_full_autotuned_and_graph():
    import torch

    from plainmp_amd import autotune
    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16")
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, THREADS)
    mv = BoxMotionValidator(pc, wl.box_width)
    E = len(wl.S1)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    with torch.cuda.stream(stream):
        s1 = torch.from_numpy(np.ascontiguousarray(wl.S1.T)).cuda()
        s2 = torch.from_numpy(np.ascontiguousarray(wl.S2.T)).cuda()
        valid = torch.empty(E, dtype=torch.uint8, device="cuda")
        nchk = torch.empty(E, dtype=torch.int32, device="cuda")
    stream.synchronize()
    schedule, _ = autotune.tune_edges(mv, pc, s1.data_ptr(), s2.data_ptr(), E, valid.data_ptr(), nchk.data_ptr(), sh,
                                      key="edges_mixed16_test", cache=None)
    mv._dev_check_motions(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
    stream.synchronize()
    assert np.array_equal(valid.cpu().numpy().astype(bool), ref.astype(bool))
    assert np.array_equal(nchk.cpu().numpy(), n_ref)
    g = mv._dev_capture_graph(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
    for rep in range(3):
        with torch.cuda.stream(stream):
            valid.fill_(7)
            nchk.fill_(-7)
        g.launch()
        stream.synchronize()
        assert np.array_equal(valid.cpu().numpy().astype(bool), ref.astype(bool)), rep
        assert np.array_equal(nchk.cpu().numpy(), n_ref), rep
    g.destroy()
    _report("edges_full", E=E, schedule=schedule.as_dict(), checked=int(n_ref.sum()), valid_frac=ref.mean())


def test_cloud_shelves_full_validity():
    wl = synth.workload("cloud_shelves")
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, THREADS).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    _report("cloud_full", B=len(ref), valid_frac=ref.mean(), rechecked=n_re)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name,B", [("plain_two_bars", 81920), ("cloud_shelves", 69632)])
def test_valid_batches_above_the_unit_pass(name, B):
    """Batches above 65,536 configurations take the lane-per-configuration
    validity kernel instead of the unit pass; validity exact."""
    base = {"plain_two_bars": 1024, "cloud_shelves": 16384}[name]
    wl = synth.workload(name, scale=B / base)
    assert len(wl.Q) == B
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, THREADS).astype(bool)
    got = pc.is_valid_batch(wl.Q)
    _report("valid_large", workload=name, B=B, valid_frac=ref.mean())
    assert np.array_equal(got, ref)


def test_grad_mixed32_full_evaluate_and_pairs():
    wl = synth.workload("grad_mixed32")
    oc, pc = oracle_cst(wl), product_cst(wl)
    assert np.array_equal(pc.is_valid_batch(wl.Q), oc.is_valid_batch(wl.Q, THREADS).astype(bool))
    vref, jref = oc.evaluate_batch(wl.Q, THREADS)
    vgot, jgot = pc.evaluate_batch(wl.Q)
    verr = np.abs(vgot - vref).max()
    pd = oc.pair_clearance_batch(wl.Q, True, THREADS)
    dist, _ = pc.pair_clearance_batch(wl.Q, False)
    derr = np.abs(dist - pd).max()
    clear = _second_gap(pd) > 1e-5
    n_rows, n_fd = rows_agree(oc, wl.Q, jgot, jref, clear, 0)
    _report("grad_full", B=len(wl.Q), val_err=verr, pair_err=derr, rows_checked=n_rows, settled_by_central_fd=n_fd)
    assert verr <= 5e-6 and derr <= 5e-6


def test_traj_dual_arm_full_rows_and_cost():
    wl = synth.workload("traj_dual_arm")
    oc, pc = oracle_cst(wl), product_cst(wl)
    off, X = wl.traj_offsets, wl.traj_X
    D = X.shape[1]
    w = np.linspace(0.5, 2.0, D)
    vals, jac, cost, grad = pc.traj_cost_grad_batch(X, list(off), list(w))
    vref, jref = oc.evaluate_batch(X, THREADS)
    verr = np.abs(vals - vref).max()
    # rows where the argmin is unambiguous: ext by the pair table, self by the
    # reference's own row being reproduced at a distance from the cutoff
    pd = oc.pair_clearance_batch(X, True, THREADS)
    clear = (_second_gap(pd) > 1e-5) & (vref[:, 0] < 0.1)
    n_rows, n_fd = rows_agree(oc, X, jac, jref, clear, 0)
    cerr = 0.0
    for i in range(len(off) - 1):
        a, b = off[i], off[i + 1]
        cref, gref = _oracle.smooth_cost(X[a:b].ravel(), b - a, D, list(w))
        cerr = max(cerr, abs(cost[i] - cref) / max(1.0, abs(cref)))
        assert np.allclose(grad[a:b].ravel(), gref, rtol=1e-5, atol=1e-6), i
    _report("traj_full", N=len(X), T=len(off) - 1, val_err=verr, ext_rows=n_rows, settled_by_central_fd=n_fd,
            cost_rel=cerr)
    assert verr <= 5e-6
    assert cerr <= 1e-5
