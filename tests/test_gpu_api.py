"""API semantics of the device path around the kernels (GPU): the scene
aliasing of set_sdf, captured edge graphs after a scene change, and calls on
alternating streams of one constraint (which share its workspaces)."""

import numpy as np
import pytest

from oracle import _oracle
from plainmp_amd import synth
from support import oracle_cst, oracle_prim, product_cst, product_sdf

pytestmark = pytest.mark.gpu


def _report(name, **kv):
    print(f"[{name}] " + " ".join(f"{k}={v}" for k, v in kv.items()), flush=True)


def _scene_cst(wl):
    """Product constraint whose primitive objects the test keeps."""
    from plainmp_amd import psdf

    cst = product_cst(synth.Workload(wl.name, wl.robot, []))
    sdfs = [product_sdf(p) for p in wl.prims]
    union = psdf.UnionSDF(sdfs)
    cst.set_sdf(union)
    return cst, union, sdfs


def test_set_sdf_aliasing():
    """set_sdf snapshots the union's membership but shares the primitive
    objects (primitive_sphere_collision.cpp:426-448): a primitive translated
    or rotated afterwards is seen by the next check, a primitive added to the
    union afterwards is not."""
    wl = synth.workload("grad_mixed32", scale=1 / 16)
    cst, union, sdfs = _scene_cst(wl)
    oprims = [oracle_prim(p) for p in wl.prims]
    oc = oracle_cst(synth.Workload(wl.name, wl.robot, []))
    oc.set_sdf(oprims)
    assert np.array_equal(cst.is_valid_batch(wl.Q), oc.is_valid_batch(wl.Q, 8).astype(bool))
    before = cst.is_valid_batch(wl.Q)
    # a box and a cylinder moved into the arm's reach, the cylinder also turned
    for i, target, yaw in [(0, (0.3, 0.0, 0.45), 0.0), (2, (0.0, 0.35, 0.6), 0.7)]:
        dx = np.array(target) - np.asarray(wl.prims[i]["position"], float)
        sdfs[i].translate(dx)
        oprims[i].translate(dx)
        if yaw:
            sdfs[i].rotate_z(yaw)
            oprims[i].rotate_z(yaw)
    oc.set_sdf(oprims)
    got = cst.is_valid_batch(wl.Q)
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    _report("aliasing_moved", B=len(ref), changed=int((got != before).sum()))
    assert (got != before).any()
    assert np.array_equal(got, ref)
    # a box added to the union after set_sdf is not part of the constraint
    big = {"type": "box", "width": [3.0, 3.0, 3.0], "position": [0.0, 0.0, 0.5], "rotation": np.eye(3).tolist()}
    union.add(product_sdf(big))
    got2 = cst.is_valid_batch(wl.Q)
    assert np.array_equal(got2, ref)
    assert got2.any()  # the added 3 m box would have made every configuration invalid


def test_graph_refuses_stale_tables():
    """A captured edge graph reads the scene tables of its capture: after a
    primitive moves, launch() raises instead of replaying stale tables; a
    graph captured again checks the moved scene."""
    import torch

    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=1 / 16)
    cst, _, sdfs = _scene_cst(wl)
    mv = BoxMotionValidator(cst, wl.box_width)
    E = len(wl.S1)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        s1 = torch.from_numpy(np.ascontiguousarray(wl.S1.T)).cuda()
        s2 = torch.from_numpy(np.ascontiguousarray(wl.S2.T)).cuda()
        valid = torch.empty(E, dtype=torch.uint8, device="cuda")
        nchk = torch.empty(E, dtype=torch.int32, device="cuda")
    stream.synchronize()
    sh = stream.cuda_stream
    g = mv._dev_capture_graph(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
    g.launch()
    stream.synchronize()
    oprims = [oracle_prim(p) for p in wl.prims]
    oc = oracle_cst(synth.Workload(wl.name, wl.robot, []))
    oc.set_sdf(oprims)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
    assert np.array_equal(valid.cpu().numpy().astype(bool), ref.astype(bool))
    dx = np.array([0.3, 0.0, 0.45]) - np.asarray(wl.prims[0]["position"], float)
    sdfs[0].translate(dx)
    oprims[0].translate(dx)
    with pytest.raises(ValueError, match="capture it again"):
        g.launch()
    g.destroy()
    g2 = mv._dev_capture_graph(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
    g2.launch()
    stream.synchronize()
    oc.set_sdf(oprims)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
    assert np.array_equal(valid.cpu().numpy().astype(bool), ref.astype(bool))
    assert np.array_equal(nchk.cpu().numpy(), n_ref)
    g2.destroy()


@pytest.mark.parametrize("name", ["plain_two_bars", "cloud_shelves"])
def test_calls_on_alternating_streams(name):
    """Validity calls of one constraint alternate between two streams with no
    host synchronisation in between; they share the context's flag and cloud
    workspaces, so each call must wait for the previous one.  Every result
    equals the oracle's."""
    import torch

    wl = synth.workload(name, scale=2048 / (1024 if name == "plain_two_bars" else 16384))
    oc, pc = oracle_cst(wl), product_cst(wl)
    rng = np.random.default_rng(5)
    batches = [wl.Q[rng.permutation(len(wl.Q))[: int(n)]] for n in (2048, 1000, 1536, 64, 2048, 777)]
    refs = [oc.is_valid_batch(Q, 8).astype(bool) for Q in batches]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ins = [torch.from_numpy(np.ascontiguousarray(Q.T)).to("cuda") for Q in batches]
    outs = [torch.empty(len(Q), dtype=torch.uint8, device="cuda") for Q in batches]
    nre = torch.zeros(len(batches), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # inputs in place; the calls below are not synchronised
    for i, (Q, q, v) in enumerate(zip(batches, ins, outs)):
        pc._dev_is_valid(q.data_ptr(), len(Q), len(Q), v.data_ptr(), nre[i:].data_ptr(), streams[i % 2].cuda_stream)
    torch.cuda.synchronize()
    for v, ref in zip(outs, refs):
        assert np.array_equal(v.cpu().numpy().astype(bool), ref)
    _report("alternating_streams", workload=name, calls=len(batches))


def test_traj_graph_replays_match_direct_call():
    """Trajectories of uneven length captured into a graph: every replay
    (outputs overwritten in between) gives the direct call's rows, costs and
    gradients, and the rows equal the oracle's evaluate."""
    import torch

    from plainmp_amd.constraint import capture_traj_graph

    wl = synth.workload("traj_dual_arm", scale=0.02)
    oc, pc = oracle_cst(wl), product_cst(wl)
    off, X = [int(o) for o in wl.traj_offsets], wl.traj_X
    N, D, C, T = len(X), pc.q_dim(), pc.cst_dim(), len(off) - 1
    w = list(np.linspace(0.5, 2.0, D))
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    with torch.cuda.stream(stream):
        x = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
        outs = [torch.empty(s, dtype=torch.float32, device="cuda") for s in ((C, N), (C, D, N), (T,), (D, N))]
    stream.synchronize()
    ptrs = [o.data_ptr() for o in outs]
    pc._dev_traj_cost_grad(x.data_ptr(), N, N, off, w, *ptrs, sh)
    stream.synchronize()
    direct = [o.cpu().numpy().copy() for o in outs]
    g = capture_traj_graph(pc, x.data_ptr(), N, N, off, w, *ptrs, sh)
    for rep in range(3):
        with torch.cuda.stream(stream):
            for o in outs:
                o.fill_(float("nan"))
        g.launch()
        stream.synchronize()
        for o, d in zip(outs, direct):
            assert np.array_equal(o.cpu().numpy(), d), rep
    g.destroy()
    vref, _ = oc.evaluate_batch(X, 8)
    assert np.abs(direct[0].T - vref).max() <= 5e-6
    _report("traj_graph", N=N, T=T)


@pytest.mark.parametrize("kind", [3, 7])
def test_ik_graph_matches_direct_iterations(kind):
    """Batched IK iterations captured into one graph (link Jacobians with the
    FP32 link pose, collision rows and the step, n_iter times) give the same
    configurations on two replays from the same start, bit for bit, and
    those of the same iterations launched one call at a time with FP64 link
    poses within 1e-4 rad."""
    import torch

    from plainmp_amd._native.ompl import _dev_capture_ik_graph

    wl = synth.workload("plain_two_bars", scale=1 / 16)
    pc = product_cst(wl)
    kin = pc.get_kin()
    link = int(kin.get_link_ids([wl.robot.sphere_specs[-1][0]])[0])
    n_nodes, D, C, R, n_iter = int(kin.n_links), pc.q_dim(), pc.cst_dim(), 96, 7
    rng = np.random.default_rng(3)
    lb, ub = np.asarray(wl.robot.lower, float), np.asarray(wl.robot.upper, float)
    Q0 = lb + (ub - lb) * rng.random((R, D))
    target = [0.3, 0.2, 0.5] + ([0.0, 0.0, 0.0, 1.0] if kind == 7 else [])
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    with torch.cuda.stream(stream):
        q = torch.from_numpy(np.ascontiguousarray(Q0.T)).cuda()
        lo = torch.from_numpy(lb).cuda()
        hi = torch.from_numpy(ub).cuda()
        tgt = torch.tensor(target, dtype=torch.float64, device="cuda")
        poses = torch.empty((n_nodes * 7, R), dtype=torch.float64, device="cuda")
        J = torch.empty((3, D, R), dtype=torch.float32, device="cuda")
        Ju = [torch.empty((3, D, R), dtype=torch.float32, device="cuda") for _ in range(3 if kind > 3 else 0)]
        vals = torch.empty((C, R), dtype=torch.float32, device="cuda")
        cjac = torch.empty((C, D, R), dtype=torch.float32, device="cuda")
    stream.synchronize()
    units = ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])
    args = (kind, link, tgt.data_ptr())
    tail = (C, lo.data_ptr(), hi.data_ptr(), 1e-3, 0.2, 0.02, 100.0)
    for _ in range(n_iter):
        pc._dev_link_poses(q.data_ptr(), R, R, 1, poses.data_ptr(), sh)
        pc._dev_link_jacobian(q.data_ptr(), R, R, link, [], J.data_ptr(), sh)
        for k in range(len(Ju)):
            pc._dev_link_jacobian(q.data_ptr(), R, R, link, units[k], Ju[k].data_ptr(), sh)
        pc._dev_evaluate(q.data_ptr(), R, R, vals.data_ptr(), cjac.data_ptr(), sh)
        pc._dev_ik_step(q.data_ptr(), R, poses.data_ptr(), *args, J.data_ptr(), [u.data_ptr() for u in Ju],
                        vals.data_ptr(), cjac.data_ptr(), *tail, sh)
    stream.synchronize()
    direct = q.cpu().numpy().copy()
    assert not np.array_equal(direct, Q0.T)  # the iterations moved the restarts
    g = _dev_capture_ik_graph(pc, q.data_ptr(), R, *args, J.data_ptr(), [u.data_ptr() for u in Ju],
                              vals.data_ptr(), cjac.data_ptr(), *tail, n_iter, sh)
    replays = []
    for rep in range(2):
        with torch.cuda.stream(stream):
            q.copy_(torch.from_numpy(np.ascontiguousarray(Q0.T)))
        g.launch()
        stream.synchronize()
        replays.append(q.cpu().numpy().copy())
    g.destroy()
    assert np.array_equal(replays[0], replays[1])
    err = np.abs(replays[0] - direct).max()
    _report("ik_graph", kind=kind, R=R, n_iter=n_iter, max_diff_vs_fp64_poses=err)
    assert err <= 1e-4


def test_graph_on_the_null_stream_orders_with_it():
    """A graph captured with no context stream (the caller's null stream)
    replays in order with the caller's null-stream work: outputs cleared on
    the null stream before each replay and read there after it hold the
    direct call's values, with no synchronization of any other stream."""
    import torch

    from plainmp_amd.constraint import capture_traj_graph

    wl = synth.workload("traj_dual_arm", scale=0.01)
    pc = product_cst(wl)
    off, X = [int(o) for o in wl.traj_offsets], wl.traj_X
    N, D, C, T = len(X), pc.q_dim(), pc.cst_dim(), len(off) - 1
    w = list(np.linspace(0.5, 2.0, D))
    x = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    outs = [torch.empty(s, dtype=torch.float32, device="cuda") for s in ((C, N), (C, D, N), (T,), (D, N))]
    ptrs = [o.data_ptr() for o in outs]
    torch.cuda.synchronize()
    pc._dev_traj_cost_grad(x.data_ptr(), N, N, off, w, *ptrs, 0)
    torch.cuda.synchronize()
    direct = [o.cpu().numpy().copy() for o in outs]
    g = capture_traj_graph(pc, x.data_ptr(), N, N, off, w, *ptrs, 0)
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        g.launch()
        for o, d in zip(outs, direct):
            assert np.array_equal(o.cpu().numpy(), d), rep
    g.destroy()
    _report("null_stream_graph", N=N, T=T)
