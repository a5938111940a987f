"""HIP path vs the FP64 oracle on the same seeded inputs.

Bars (stated per test):
* validity booleans: bit-exact wherever the oracle's clearance is farther than
  the 1e-5 m band from zero; inside the band the GPU re-decides in FP64, and
  the tests require exactness there too;
* FP32 sphere centres / clearances: max |FP32 - FP64| <= 5e-6 m;
* Jacobians: relative error <= 1e-4 against FP64 central differences.
"""

import numpy as np
import pytest

from oracle import _oracle
from plainmp_amd import synth
from support import oracle_cst, oracle_min_margin, product_cst

pytestmark = pytest.mark.gpu

BAND = 1e-5


def _report(name, **kv):
    print(f"[{name}] " + " ".join(f"{k}={v}" for k, v in kv.items()), flush=True)


def test_valid_plain_two_bars_full():
    wl = synth.workload("plain_two_bars")  # 1,024 configs, seed 0
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    margin = oracle_min_margin(oc, wl.Q)
    outside = margin > BAND
    _report("valid_plain", B=len(ref), valid_frac=ref.mean(), rechecked=n_re, in_band=int((~outside).sum()))
    assert np.array_equal(got[outside], ref[outside])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n_configs", [16384, 20480])
def test_valid_both_batch_paths(n_configs):
    """Cloud-free validity takes a lane per (configuration, unit) up to 16,384
    configurations and a lane per configuration above: both answers equal
    the oracle's, bit-exact (the band re-decided in FP64)."""
    wl = synth.workload("plain_two_bars", scale=n_configs / 1024)
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    _report("valid_paths", B=len(ref), valid_frac=ref.mean(), rechecked=n_re)
    assert len(ref) == n_configs
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("bad", [np.nan, np.inf, 1e30, 20.0])
@pytest.mark.parametrize("n_configs", [1024, 20480, 81920])
def test_valid_nan_configurations(n_configs, bad):
    """A NaN joint value makes every SDF comparison false, so the reference
    reports the configuration invalid (AABB pre-test not outside, is_outside
    false); an infinite one too.  A huge finite one (1e30 rad) overflows the
    FP32 walk to NaN but not the reference's FP64 one, which finds the robot
    far away: the FP32 NaN goes to the FP64 walk.  Both the unit pass (up to
    65,536 configurations) and the one-lane pass agree with the oracle."""
    wl = synth.workload("plain_two_bars", scale=n_configs / 1024)
    Q = wl.Q.copy()
    rng = np.random.default_rng(4)
    rows = rng.choice(len(Q), 37, replace=False)
    Q[rows, rng.integers(0, Q.shape[1], size=37)] = bad
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(Q)
    _report("valid_bad_values", B=len(Q), bad=bad, rows=len(rows), ref_valid_rows=int(ref[rows].sum()),
            rechecked=n_re)
    if not np.isfinite(bad):
        assert not ref[rows].any()
    assert np.array_equal(got, ref)


def test_edges_beyond_the_fp32_ranges():
    """Edges whose endpoints hold joint values beyond the FP32 walk's ranges
    (angles past 4 pi) are walked in FP64 state by state: validity and the
    reference-order checked counts match the oracle (motion_validator.cpp:17-79)."""
    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=1 / 16)  # 256 edges
    S1, S2 = wl.S1.copy(), wl.S2.copy()
    far = np.arange(0, len(S1), 3)  # every third edge: one joint shifted by 8 turns at both ends
    S1[far, 2] += 16 * np.pi
    S2[far, 2] += 16 * np.pi
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref, n_ref = oc.check_motions(S1, S2, True, list(wl.box_width), 0.0, 8)
    got, n_got = BoxMotionValidator(pc, wl.box_width).check_motions(S1, S2)
    _report("edges_far", E=len(S1), far=len(far), valid_far=float(ref[far].mean()))
    assert np.array_equal(got, ref.astype(bool))
    assert np.array_equal(n_got, n_ref)


def test_clearance_and_centers_plain_two_bars():
    wl = synth.workload("plain_two_bars")
    oc, pc = oracle_cst(wl), product_cst(wl)
    Q = wl.Q[:256]
    for accurate in (False, True):
        ref = oc.sphere_clearance_batch(Q, accurate, 8)
        got = pc.sphere_clearance_batch(Q, accurate)
        err = np.abs(got - ref).max()
        cref = oc.sphere_centers_batch(Q, accurate)
        cgot = pc.sphere_centers_batch(Q, accurate)
        cerr = np.abs(cgot - cref).max()
        _report("clearance", accurate=accurate, max_err=err, center_err=cerr)
        assert err <= 5e-6
        assert cerr <= 5e-6


def test_edges_mixed16_subset():
    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=0.125)  # 512 edges
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
    mv = BoxMotionValidator(pc, wl.box_width)
    got, n_got = mv.check_motions(wl.S1, wl.S2)
    _report("edges", E=len(ref), valid_frac=ref.mean(), mean_checked_ref=n_ref.mean(), mean_checked_gpu=n_got.mean(),
            rechecked=mv.last_recheck_count())
    assert np.array_equal(got, ref.astype(bool))
    # states evaluated in the reference's visiting order, up to the first invalid one
    assert np.array_equal(n_got, n_ref)
    # single-edge entry point agrees with the batch
    for e in range(8):
        assert mv.checkMotion(wl.S1[e], wl.S2[e]) == bool(ref[e])


@pytest.mark.parametrize("band", [None, 0.002, 0.05])
def test_edges_graph_replays(band):
    """An edge batch captured as a HIP graph and replayed several times, with
    direct launches of a second batch on the same context and stream in
    between: every replay and every direct call equals the oracle.  The
    kernels reset their own work counters and pending list at the end of a
    launch, so nothing carries over from one launch to the next.  A 2 mm band
    queues states whose uncertainty is one candidate (re-decided on that
    candidate alone, with and without a collision among them); the 5 cm band
    sends many edges through the full FP64 walk."""
    import torch

    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=1 / 16)  # 256 edges
    oc, pc = oracle_cst(wl), product_cst(wl)
    if band is not None:
        pc.set_band(band)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
    E, E2 = len(ref), 96
    mv = BoxMotionValidator(pc, wl.box_width)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def soa(x):
        return torch.from_numpy(np.ascontiguousarray(x.T)).to("cuda")

    with torch.cuda.stream(stream):
        s1, s2 = soa(wl.S1), soa(wl.S2)
        t1, t2 = soa(wl.S1[-E2:]), soa(wl.S2[-E2:])
        valid = torch.full((E,), 7, dtype=torch.uint8, device="cuda")
        nchk = torch.full((E,), -5, dtype=torch.int32, device="cuda")
        valid2 = torch.empty(E2, dtype=torch.uint8, device="cuda")
        nchk2 = torch.empty(E2, dtype=torch.int32, device="cuda")
        graph = mv._dev_capture_graph(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
    stream.synchronize()
    for rep in range(4):
        with torch.cuda.stream(stream):
            valid.fill_(7)
            nchk.fill_(-5)
            graph.launch()
        stream.synchronize()
        got, n_got = valid.cpu().numpy(), nchk.cpu().numpy()
        _report("edges_graph", band=band, replay=rep, valid_frac=got.astype(bool).mean())
        assert np.array_equal(got, ref.astype(np.uint8))
        assert np.array_equal(n_got, n_ref)
        with torch.cuda.stream(stream):
            mv._dev_check_motions(t1.data_ptr(), t2.data_ptr(), E2, E2, valid2.data_ptr(), nchk2.data_ptr(), sh)
        stream.synchronize()
        assert np.array_equal(valid2.cpu().numpy(), ref[-E2:].astype(np.uint8))
        assert np.array_equal(nchk2.cpu().numpy(), n_ref[-E2:])
    del graph


def test_check_path_and_shortcut():
    """Batched path utilities over the edge pass, on a detour: a valid path
    from a to b through a via state, where the direct motion a -> b collides.
    check_path's per-segment validity equals the oracle's checkMotion on
    consecutive states; the batched shortcut keeps the end states, returns a
    subsequence of the input no longer than it that still needs an
    intermediate state, every segment it adds is a valid motion per the
    oracle, and it is reproducible from its seed."""
    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=1 / 16)
    oc, pc = oracle_cst(wl), product_cst(wl)
    w = list(wl.box_width)
    ev, _ = oc.check_motions(wl.S1, wl.S2, True, w, 0.0, 8)
    e = int(np.flatnonzero(~ev.astype(bool))[0])
    a, b = wl.S1[e], wl.S2[e]
    V = np.concatenate([wl.S1, wl.S2])
    V = V[oc.is_valid_batch(V, 8).astype(bool)]
    r1, _ = oc.check_motions(np.repeat(a[None], len(V), 0), V, True, w, 0.0, 8)
    r2, _ = oc.check_motions(V, np.repeat(b[None], len(V), 0), True, w, 0.0, 8)
    v = V[np.flatnonzero(r1.astype(bool) & r2.astype(bool))[0]]
    t = np.linspace(0, 1, 21)[:, None]
    P = np.concatenate([a + t * (v - a), (v + t * (b - v))[1:]])
    P[1:-1] += np.random.default_rng(7).normal(scale=0.01, size=P[1:-1].shape)
    mv = BoxMotionValidator(pc, w)
    seg = mv.check_path(P)
    ref, _ = oc.check_motions(P[:-1], P[1:], True, w, 0.0, 8)
    assert np.array_equal(seg, ref.astype(bool))
    S = mv.shortcut_path(P, 256, 8, 3)
    idx = [int(np.flatnonzero((P == s).all(axis=1))[0]) for s in S]
    assert idx[0] == 0 and idx[-1] == len(P) - 1 and all(x < y for x, y in zip(idx, idx[1:]))
    new = [(x, y) for x, y in zip(idx, idx[1:]) if y > x + 1]
    _report("shortcut", states_in=len(P), states_out=len(S), spliced=len(new), valid_segments_in=int(seg.sum()))
    assert len(S) >= 3 and len(new) > 0
    r, _ = oc.check_motions(P[[x for x, _ in new]], P[[y for _, y in new]], True, w, 0.0, 8)
    assert r.all()

    def length(X):
        return np.linalg.norm(np.diff(X, axis=0), axis=1).sum()

    assert length(S) <= length(P) + 1e-12
    assert np.array_equal(S, mv.shortcut_path(P, 256, 8, 3))
    assert len(mv.check_path(P[:1])) == 0


def test_edges_euclidean_subset():
    from plainmp_amd.ompl import EuclideanMotionValidator

    wl = synth.workload("edges_mixed16", scale=0.0625)
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, False, [], 0.1, 8)
    got, n_got = EuclideanMotionValidator(pc, 0.1).check_motions(wl.S1, wl.S2)
    assert np.array_equal(got, ref.astype(bool))
    assert np.array_equal(n_got, n_ref)


def _second_gap(d):
    """Gap between the two smallest entries of each row (ties make argmins ambiguous)."""
    s = np.sort(d.reshape(len(d), -1), axis=1)
    return s[:, 1] - s[:, 0]


def test_evaluate_rows_grad_mixed32():
    wl = synth.workload("grad_mixed32", scale=1 / 32)  # 256 configs
    oc, pc = oracle_cst(wl), product_cst(wl)
    vref, jref = oc.evaluate_batch(wl.Q, 8)
    vgot, jgot = pc.evaluate_batch(wl.Q)
    assert vgot.shape == vref.shape and jgot.shape == jref.shape
    verr = np.abs(vgot - vref).max()
    pd = oc.pair_clearance_batch(wl.Q, True, 8)
    clear = _second_gap(pd) > 1e-5  # unambiguous argmin pair
    jerr = np.abs(jgot - jref)[clear].max()
    scale = np.maximum(np.abs(jref[clear]), 1.0)
    _report("evaluate", B=len(vref), val_err=verr, jac_err=jerr, ambiguous=int((~clear).sum()))
    assert verr <= 5e-6
    # reference rows use a forward-difference SDF gradient (eps 1e-6): compare at 1e-4
    assert (np.abs(jgot - jref)[clear] <= 1e-4 * scale).all()


@pytest.mark.parametrize("name", ["grad_mixed32", "cloud_shelves"])
def test_pair_clearance_grad_vs_central_fd(name):
    """Every sphere-primitive distance against the oracle and its gradient
    against FP64 central differences (32 mixed primitives; the 65,536-point
    shelf cloud, whose distance is sqrt of the nearest squared distance minus
    r_cloud and whose gradient is taken at the nearest point)."""
    wl = synth.workload(name, scale=64 / (16384 if name == "cloud_shelves" else 8192))  # 64 configs
    oc, pc = oracle_cst(wl), product_cst(wl)
    dref = oc.pair_clearance_batch(wl.Q, True, 8)
    dist, grad = pc.pair_clearance_batch(wl.Q, True)
    assert np.abs(dist - dref).max() <= 5e-6
    h = 1e-6
    B, S, P, D = grad.shape
    fd = np.empty((B, S, P, D))
    fd2 = np.empty((B, S, P, D))
    for j in range(D):
        E = np.zeros(D)
        E[j] = 1.0
        fd[..., j] = (oc.pair_clearance_batch(wl.Q + h * E, True, 8) - oc.pair_clearance_batch(wl.Q - h * E, True, 8)) / (2 * h)
        fd2[..., j] = (oc.pair_clearance_batch(wl.Q + 4 * h * E, True, 8) - oc.pair_clearance_batch(wl.Q - 4 * h * E, True, 8)) / (8 * h)
    smooth = np.abs(fd - fd2) <= 1e-6  # away from SDF kinks
    rel = np.abs(grad - fd) / np.maximum(np.abs(fd), 1e-2)
    _report("pair_grad", entries=fd.size, smooth=int(smooth.sum()), max_rel=rel[smooth].max())
    assert rel[smooth].max() <= 1e-4


def test_self_collision_dual_arm():
    wl = synth.workload("traj_dual_arm", scale=0.005)  # ~10 trajectories of waypoints
    oc, pc = oracle_cst(wl), product_cst(wl)
    Q = wl.traj_X[:512]
    ref = oc.is_valid_batch(Q, 8).astype(bool)
    got = pc.is_valid_batch(Q)
    _report("dual_arm", B=len(Q), valid_frac=ref.mean())
    assert np.array_equal(got, ref)
    vref, jref = oc.evaluate_batch(Q[:64], 8)
    vgot, jgot = pc.evaluate_batch(Q[:64])
    assert np.abs(vgot - vref).max() <= 5e-6


@pytest.mark.parametrize("base", ["PLANAR", "FLOATING"])
def test_base_types(base):
    from plainmp_amd.kinematics import BaseType

    wl = synth.workload("grad_mixed32", scale=1 / 64)
    extra = 3 if base == "PLANAR" else 6
    rng = np.random.default_rng(11)
    Q = np.concatenate([wl.Q, rng.normal(0, 0.1, size=(len(wl.Q), extra))], axis=1)
    oc = oracle_cst(wl, base_type=getattr(_oracle.BaseType, base))
    pc = product_cst(wl, base_type=getattr(BaseType, base))
    assert np.array_equal(pc.is_valid_batch(Q), oc.is_valid_batch(Q, 8).astype(bool))
    vref, jref = oc.evaluate_batch(Q, 8)
    vgot, jgot = pc.evaluate_batch(Q)
    assert np.abs(vgot - vref).max() <= 5e-6
    pd = oc.pair_clearance_batch(Q, True, 8)
    clear = _second_gap(pd) > 1e-5
    # base columns of the reference are forward differences with eps 1e-7
    assert (np.abs(jgot - jref)[clear] <= 2e-4 * np.maximum(np.abs(jref[clear]), 1.0)).all()


def test_single_config_api_matches_batch():
    wl = synth.workload("plain_two_bars", scale=1 / 64)
    oc, pc = oracle_cst(wl), product_cst(wl)
    for q in wl.Q[:4]:
        assert pc.is_valid(q) == oc.is_valid(q)
        v, j = pc.evaluate(q)
        vr, jr = oc.evaluate(q)
        assert np.allclose(v, vr, atol=5e-6)
        pc.update_kintree(q, False)
        oc.update_kintree(q, False)
        got = np.array([c for c, _ in pc.get_all_spheres()])
        ref, _ = oc.get_all_spheres()
        assert np.abs(got - ref).max() <= 5e-6
        gs = np.array([c for c, _ in pc.get_group_spheres()])
        gref, _ = oc.get_group_spheres()
        assert np.abs(gs - gref).max() <= 5e-6


def test_psdf_against_oracle():
    from plainmp_amd import psdf
    from support import oracle_prim, product_sdf

    prims = synth.mixed_primitives(12, seed=7) + [{"type": "ground", "height": 0.3}]
    rng = np.random.default_rng(3)
    pts = rng.normal(0.5, 0.6, size=(3, 500))
    for p in prims:
        o, g = oracle_prim(p), product_sdf(p)
        assert np.allclose(g.evaluate_batch(pts), o.evaluate_batch(pts), atol=1e-12, rtol=0)
        for r in (0.0, 0.05, 0.3):
            got = g.is_outside_batch(pts, r)
            ref = np.array([o.is_outside(pts[:, i], r) for i in range(pts.shape[1])])
            assert np.array_equal(got, ref)
        assert np.allclose(g.lb, o.lb) and np.allclose(g.ub, o.ub)
    u = psdf.UnionSDF([product_sdf(p) for p in prims[:4]])
    ref = np.min([oracle_prim(p).evaluate_batch(pts) for p in prims[:4]], axis=0)
    assert np.allclose(u.evaluate_batch(pts), ref, atol=1e-12, rtol=0)
    assert abs(psdf.GroundSDF(1.0).evaluate(np.zeros(3)) - 1.0) == 0.0  # tests/test_psdf.py:201-204


def test_kdtree_and_cloud():
    from plainmp_amd import psdf
    from plainmp_amd.kdtree import KDTree

    rng = np.random.default_rng(5)
    pts = rng.normal(size=(1000, 3))
    tree = KDTree(pts)
    for _ in range(50):
        q = rng.normal(size=3)
        i = np.argmin(((pts - q) ** 2).sum(1))
        assert np.allclose(tree.query(q), pts[i])
        assert np.isclose(tree.sqdist(q), ((pts[i] - q) ** 2).sum())
    obst = rng.uniform(size=(100, 3))
    cs = psdf.CloudSDF(obst, 0.1)
    ev = rng.normal(size=(3, 300)) * 1.5
    ref = _oracle.Prim.cloud(obst, 0.1).evaluate_batch(ev)
    assert np.allclose(cs.evaluate_batch(ev), ref, atol=1e-12)


def _tile_fields(d):
    """The field bytes of the tile tables (the 8 padding bytes of each
    48-byte FP32 tile record carry no value)."""
    d = dict(d)
    d["tiles"] = np.frombuffer(d["tiles"], np.uint8).reshape(-1, 48)[:, :40].tobytes()
    return d


@pytest.mark.parametrize("case", ["one", "ragged65", "flat_axis", "duplicates", "shelves65536", "wide1M"])
def test_cloud_tile_build_matches_oracle(case):
    """The device build of a cloud's tile tables (bounds, Morton sort, tiles,
    matrix-core operands, super-tiles; pmp_cloud_build_tiles) against the
    oracle's restatement, bit-exact in every table."""
    from plainmp_amd import kdtree as kd

    rng = np.random.default_rng(11)
    if case == "one":
        P = np.array([[0.1, -0.2, 0.3]])
    elif case == "ragged65":
        P = rng.uniform(-1, 1, (65, 3))
    elif case == "flat_axis":  # zero span on z: every code has z = 0
        P = rng.uniform(-1, 1, (1000, 3))
        P[:, 2] = 0.25
    elif case == "duplicates":  # equal codes, ordered by index
        P = rng.uniform(-1, 1, (3000, 3))
        P[100:900] = P[5]
    elif case == "shelves65536":
        P = synth.shelf_cloud(65536, seed=2)
    else:
        P = rng.normal(size=(1 << 20, 3)) * np.array([2.0, 1.0, 0.5])
    got = _tile_fields(kd.build_cloud_tiles(P))
    ref = _tile_fields(_oracle.cloud_tiles(P))
    n_tiles, n_supers, _ = kd.cloud_tile_sizes(len(P))
    _report("cloud_tiles", case=case, n=len(P), tiles=n_tiles, supers=n_supers)
    assert len(ref["tiles"]) == 40 * n_tiles and len(ref["supers"]) == 32 * n_supers
    for k in ref:
        assert got[k] == ref[k], k


@pytest.mark.parametrize("n_configs", [256])
def test_cloud_valid_and_clearance(n_configs):
    wl = synth.workload("cloud_shelves", scale=n_configs / 16384)
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    _report("cloud_valid", B=len(ref), valid_frac=ref.mean(), rechecked=n_re)
    assert np.array_equal(got, ref)
    Q = wl.Q[: min(16, len(wl.Q))]
    cref = oc.sphere_clearance_batch(Q, False, 8)
    cgot = pc.sphere_clearance_batch(Q, False)
    err = np.abs(cgot - cref).max()
    _report("cloud_clearance", max_err=err)
    assert err <= 5e-6


@pytest.mark.parametrize("band", [None, 0.02])
def test_edges_cloud(band):
    """Motion checks against the 65,536-point shelf cloud (the reference's
    planning-with-a-cloud use): the edge pass tests the cloud inside its walk
    and sends in-band states to the FP64 walk; validity and the states
    checked in the reference's visiting order equal the oracle's KD-tree
    answers.  The 2 cm band forces FP64 re-decisions of cloud states."""
    from plainmp_amd.ompl import BoxMotionValidator

    E = 256
    wl = synth.workload("cloud_shelves", scale=2 * E / 16384)
    oc, pc = oracle_cst(wl), product_cst(wl)
    if band is not None:
        pc.set_band(band)
    S1, S2 = wl.Q[:E], wl.Q[E : 2 * E]
    w = [0.05] * S1.shape[1]
    ref, n_ref = oc.check_motions(S1, S2, True, w, 0.0, 8)
    mv = BoxMotionValidator(pc, w)
    got, n_got = mv.check_motions(S1, S2)
    _report("edges_cloud", band=band, E=E, valid_frac=ref.mean(), rechecked=mv.last_recheck_count())
    if band is not None:
        assert mv.last_recheck_count() > 0
    assert np.array_equal(got, ref.astype(bool))
    assert np.array_equal(n_got, n_ref)


def _sparse_cloud_workload(n_configs, seed=21):
    """The 8-DOF mobile torso arm against 1,000 points spread uniformly over a
    4 m cube around it: 64-point tiles span metres, where the FP32 tile-local
    squared distance carries an error far above the 1e-5 m band."""
    import dataclasses

    rng = np.random.default_rng(seed)
    base = synth.workload("cloud_shelves", scale=n_configs / 16384)
    cloud = rng.uniform([-2.0, -2.0, -1.4], [2.0, 2.0, 2.6], size=(1000, 3))
    return dataclasses.replace(base, cloud=cloud)


def test_sparse_cloud_valid_edges_and_clearance():
    """Validity, motion checks and clearances against a sparse cloud: the
    validity decisions bracket the FP32 distance by the per-tile error bound
    (|c'|^2 + rho^2 term), so the booleans stay exact; the clearance error is
    the FP32 GEMM form's (stated bound 2e-6 (|c'|^2 + rho^2) / (2 d))."""
    from plainmp_amd import kdtree as kd
    from plainmp_amd.ompl import BoxMotionValidator

    wl = _sparse_cloud_workload(4096)
    oc, pc = oracle_cst(wl), product_cst(wl)
    sizes = kd.cloud_tile_sizes(len(wl.cloud))
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    _report("sparse_cloud_valid", B=len(ref), tiles=sizes[0], valid_frac=ref.mean(), rechecked=n_re)
    assert 0.05 < ref.mean() < 0.95
    assert np.array_equal(got, ref)
    E = 512
    S1, S2 = wl.Q[:E], wl.Q[E: 2 * E]
    w = [0.05] * S1.shape[1]
    eref, n_ref = oc.check_motions(S1, S2, True, w, 0.0, 8)
    mv = BoxMotionValidator(pc, w)
    egot, n_got = mv.check_motions(S1, S2)
    _report("sparse_cloud_edges", E=E, valid_frac=eref.mean(), rechecked=mv.last_recheck_count())
    assert np.array_equal(egot, eref.astype(bool))
    assert np.array_equal(n_got, n_ref)
    Q = wl.Q[:64]
    cerr = np.abs(pc.sphere_clearance_batch(Q, False) - oc.sphere_clearance_batch(Q, False, 8)).max()
    _report("sparse_cloud_clearance", max_err=cerr)
    assert cerr <= 5e-5


@pytest.mark.parametrize("height", [0.0, 0.3])
def test_valid_every_kind_with_ground(height):
    """Boxes (axis-aligned and rotated), cylinders (z-aligned and rotated),
    spheres and a ground plane in one scene: the kind-sorted validity pass
    against the oracle, and the edge kernel on the same scene."""
    import dataclasses

    from plainmp_amd.ompl import BoxMotionValidator

    base = synth.workload("grad_mixed32", scale=1 / 16)  # 512 configs, 32 mixed primitives
    wl = dataclasses.replace(base, prims=base.prims + [{"type": "ground", "height": height}])
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref = oc.is_valid_batch(wl.Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(wl.Q)
    _report("every_kind", height=height, B=len(ref), valid_frac=ref.mean(), rechecked=n_re)
    assert np.array_equal(got, ref)
    S1, S2 = wl.Q[:128], wl.Q[128:256]
    width = np.full(S1.shape[1], 0.1)
    eref, _ = oc.check_motions(S1, S2, True, list(width), 0.0, 8)
    egot, _ = BoxMotionValidator(pc, width).check_motions(S1, S2)
    assert np.array_equal(egot, eref.astype(bool))


@pytest.mark.parametrize("name", ["plain_two_bars", "every_kind", "dual_arm", "cloud", "edges"])
def test_fp64_redecision_wide_band(name):
    """A 5 cm band sends a large share of the work to the FP64 pass (block
    visitor with the cull table, multi-wave edge walk with split candidate
    lists); its decisions must still equal the oracle's."""
    import dataclasses

    from plainmp_amd.ompl import BoxMotionValidator

    if name == "every_kind":
        base = synth.workload("grad_mixed32", scale=1 / 32)
        wl = dataclasses.replace(base, prims=base.prims + [{"type": "ground", "height": 0.0}])
        Q = wl.Q
    elif name == "dual_arm":
        wl = synth.workload("traj_dual_arm", scale=0.005)
        Q = wl.traj_X[:256]
    elif name == "cloud":
        wl = synth.workload("cloud_shelves", scale=64 / 16384)
        Q = wl.Q
    elif name == "edges":
        wl = synth.workload("edges_mixed16", scale=1 / 16)
    else:
        wl = synth.workload(name, scale=0.25)
        Q = wl.Q
    oc, pc = oracle_cst(wl), product_cst(wl)
    pc.set_band(0.05)
    if name == "edges":
        ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
        mv = BoxMotionValidator(pc, wl.box_width)
        got, n_got = mv.check_motions(wl.S1, wl.S2)
        _report("wide_band_edges", E=len(ref), rechecked=mv.last_recheck_count())
        assert mv.last_recheck_count() > 0
        assert np.array_equal(got, ref.astype(bool))
        assert np.array_equal(n_got, n_ref)
        return
    ref = oc.is_valid_batch(Q, 8).astype(bool)
    got, n_re = pc.is_valid_batch_stats(Q)
    _report("wide_band_" + name, B=len(ref), rechecked=n_re, valid_frac=ref.mean())
    assert n_re > 0
    assert np.array_equal(got, ref)


def test_trajectory_rows_and_smoothness_dual_arm():
    """One launch over uneven trajectories: per-waypoint evaluate rows against
    the oracle's evaluate, smoothness cost and gradient against the oracle's
    smooth_cost (x^T P x and 2 P x of smoothcost_fullmat)."""
    wl = synth.workload("traj_dual_arm", scale=0.01)  # ~20 trajectories, 20-100 waypoints each
    oc, pc = oracle_cst(wl), product_cst(wl)
    off, X = wl.traj_offsets, wl.traj_X
    D = X.shape[1]
    w = np.linspace(0.5, 2.0, D)
    vals, jac, cost, grad = pc.traj_cost_grad_batch(X, list(off), list(w))
    vref, jref = oc.evaluate_batch(X, 8)
    assert vals.shape == vref.shape and jac.shape == jref.shape
    assert np.abs(vals - vref).max() <= 5e-6
    pd = oc.pair_clearance_batch(X, True, 8)
    clear = (_second_gap(pd) > 1e-5) & (vref[:, 0] < 0.1)  # unambiguous ext argmin below the cutoff
    assert clear.sum() > 0
    scale = np.maximum(np.abs(jref[clear, 0]), 1.0)
    assert (np.abs(jac[clear, 0] - jref[clear, 0]) <= 1e-4 * scale).all()
    for i in range(len(off) - 1):
        a, b = off[i], off[i + 1]
        cref, gref = _oracle.smooth_cost(X[a:b].ravel(), b - a, D, list(w))
        assert abs(cost[i] - cref) <= 1e-5 * max(1.0, abs(cref))
        assert np.allclose(grad[a:b].ravel(), gref, rtol=1e-5, atol=1e-6)
    _report("traj", N=len(X), T=len(off) - 1, val_err=np.abs(vals - vref).max())


def test_evaluate_and_trajectory_cloud():
    """evaluate rows and trajectory rows against the 65,536-point shelf cloud:
    values (sqrt of the nearest squared distance - r_cloud - r, below the
    0.1 m cutoff) and rows (gradient at the nearest point; the reference's
    forward difference approximates it) against the oracle's KD-tree
    evaluate, and the trajectory launch against the plain one."""
    wl = synth.workload("cloud_shelves", scale=256 / 16384)
    oc, pc = oracle_cst(wl), product_cst(wl)
    Q = wl.Q
    vref, jref = oc.evaluate_batch(Q, 8)
    vgot, jgot = pc.evaluate_batch(Q)
    assert vgot.shape == vref.shape and jgot.shape == jref.shape
    verr = np.abs(vgot - vref).max()
    pd = oc.pair_clearance_batch(Q, True, 8)
    clear = (_second_gap(pd) > 1e-5) & (vref[:, 0] < 0.1)
    _report("evaluate_cloud", B=len(Q), below_cutoff=int((vref[:, 0] < 0.1).sum()), clear=int(clear.sum()), val_err=verr)
    assert clear.sum() >= 10
    assert verr <= 5e-6
    scale = np.maximum(np.abs(jref[clear]), 1.0)
    assert (np.abs(jgot - jref)[clear] <= 1e-4 * scale).all()
    off = [0, 40, 100, 101, 256]
    w = np.linspace(0.5, 2.0, Q.shape[1])
    vals, jac, cost, grad = pc.traj_cost_grad_batch(Q, off, list(w))
    assert np.abs(vals - vref).max() <= 5e-6
    assert (np.abs(jac - jref)[clear] <= 1e-4 * scale).all()
    for i in range(len(off) - 1):
        a, b = off[i], off[i + 1]
        cref, gref = _oracle.smooth_cost(Q[a:b].ravel(), b - a, Q.shape[1], list(w))
        assert abs(cost[i] - cref) <= 1e-5 * max(1.0, abs(cref))


@pytest.mark.parametrize("with_box", [True, False])
def test_sqp_ineq_csc_dual_arm(with_box):
    """The SQP inequality sequence (collision rows of every waypoint + the
    motion-step box) of many trajectories, assembled on the GPU as CSC, against
    the oracle's SequentialCst restatement: the sparsity structure (column
    pointers and row indices) exactly, box rows and entries exactly, collision
    values within 5e-6 and collision entries within 1e-4 (relative) where the
    waypoint's argmin pair is unambiguous.  Empty and one-waypoint
    trajectories are included."""
    from plainmp_amd.constraint import sqp_ineq_matrices

    wl = synth.workload("traj_dual_arm", scale=0.004)
    oc, pc = oracle_cst(wl), product_cst(wl)
    off0, X0 = list(wl.traj_offsets), wl.traj_X
    X = np.concatenate([X0, X0[:1]])
    off = [0] + [o for o in off0[1:]] + [off0[-1], off0[-1] + 1]  # then an empty and a one-waypoint trajectory
    D = X.shape[1]
    msbox = list(np.linspace(0.05, 0.2, D)) if with_box else None
    got = sqp_ineq_matrices(pc, X, off, msbox)
    pd = oc.pair_clearance_batch(X, True, 8)
    clear_wp = _second_gap(pd) > 1e-5
    C = pc.cst_dim()
    assert len(got) == len(off) - 1
    for i, (c, J) in enumerate(got):
        a, b = off[i], off[i + 1]
        T = b - a
        cr, cpr, rir, var, (rows, cols) = _oracle.seq_ineq_evaluate(oc, X[a:b].reshape(T, D), msbox)
        assert J.shape == (rows, cols) and len(c) == rows
        assert np.array_equal(J.indptr, cpr) and np.array_equal(J.indices, rir)
        nc = C * T
        assert np.abs(c[:nc] - cr[:nc]).max(initial=0) <= 5e-6
        assert np.array_equal(c[nc:], cr[nc:])
        box = J.indices >= nc
        assert np.array_equal(J.data[box], var[box])
        col_wp = np.repeat(np.arange(T), D)  # waypoint of each column
        entry_wp = np.repeat(col_wp, np.diff(J.indptr))
        sel = (~box) & clear_wp[a + entry_wp]
        assert (np.abs(J.data[sel] - var[sel]) <= 1e-4 * np.maximum(np.abs(var[sel]), 1.0)).all()
    _report("sqp_csc", box=with_box, trajectories=len(got), waypoints=len(X), nnz=sum(J.nnz for _, J in got))


def test_robot_spec_yaml_ingestion(tmp_path):
    """A robot described by a plainmp YAML configuration (sphere groups in an
    included file, only-self links, self pairs given as pairs and as group
    pairs, a base-fixed box with a yaw) gives, through RobotSpec, a
    constraint whose validity and evaluate values equal the oracle's for the
    same spheres, pairs and fixed primitive."""
    import yaml

    from plainmp_amd import psdf
    from plainmp_amd.robot_spec import RobotSpec, rpy_matrix
    from plainmp_amd.synth import build_sdf

    wl = synth.workload("traj_dual_arm", scale=0.004)
    r = wl.robot
    spheres = {"collision_spheres": {
        ln: {"spheres": [[float(v) for v in c] + [float(rr)] for c, rr in zip(pos, rad)]}
        for ln, pos, rad, _ in r.sphere_specs}}
    pairs = [list(p) for p in r.self_pairs]
    # placed where it changes answers, and where the opposite yaw would change others
    box = {"type": "box", "extents": [0.3, 0.8, 0.4], "position": [0.35, 0.0, 0.4], "rotation": [0.0, 0.0, 0.4]}
    conf = {"control_joint_names": list(r.joint_names),
            "only_self_collision_links": [ln for ln, _, _, only in r.sphere_specs if only],
            "self_collision_pairs": pairs[:2],
            "self_collision_group_pairs": [{"group1": [a], "group2": [b]} for a, b in pairs[2:]],
            "self_body_collision_primitives": [box],
            "include": ["spheres.yaml"]}
    (tmp_path / "spheres.yaml").write_text(yaml.safe_dump(spheres))
    (tmp_path / "robot.yaml").write_text(yaml.safe_dump(conf))
    cst = RobotSpec(tmp_path / "robot.yaml", urdf=r.urdf).create_collision_const()
    cst.set_sdf(psdf.UnionSDF([build_sdf(p) for p in wl.prims]))
    fixed = {"type": "box", "width": box["extents"], "position": box["position"], "rotation": rpy_matrix(0.4, 0.0, 0.0)}
    oc = oracle_cst(wl, fixed_prims=[fixed])
    Q = wl.traj_X[:256]
    ref = oc.is_valid_batch(Q, 8).astype(bool)
    got = cst.is_valid_batch(Q)
    vref, _ = oc.evaluate_batch(Q, 8)
    vgot, _ = cst.evaluate_batch(Q)
    _report("robot_spec", B=len(Q), valid_frac=ref.mean(), val_err=np.abs(vgot - vref).max())
    assert 0.0 < ref.mean() < 1.0
    assert (oracle_cst(wl).is_valid_batch(Q, 8).astype(bool) != ref).any()  # the box matters
    assert np.array_equal(got, ref)
    assert np.abs(vgot - vref).max() <= 5e-6


def test_trajectory_edge_cases():
    """Empty, one- and two-waypoint trajectories have zero smoothness cost and
    gradient; bad offsets are rejected as ValueError."""
    wl = synth.workload("grad_mixed32", scale=1 / 256)
    pc = product_cst(wl)
    X = wl.Q[:7]
    off = [0, 0, 1, 3, 3, 7]
    vals, jac, cost, grad = pc.traj_cost_grad_batch(X, off)
    assert cost[0] == 0 and cost[1] == 0 and cost[2] == 0 and cost[3] == 0
    assert np.all(grad[:3] == 0)
    cref, gref = _oracle.smooth_cost(X[3:7].ravel(), 4, X.shape[1], [1.0] * X.shape[1])
    assert abs(cost[4] - cref) <= 1e-5 * max(1.0, cref)
    with pytest.raises(ValueError):
        pc.traj_cost_grad_batch(X, [0, 3, 2, 7])
    with pytest.raises(ValueError):
        pc.traj_cost_grad_batch(X, [0, 3, 6])


@pytest.mark.parametrize("whole_chunks", [1, 0])
def test_evaluate_unit_modes(whole_chunks):
    """Evaluate rows and trajectories with each configuration walked as one
    unit (whole_chunks = 1) and split per sphere run and self pair (the
    default for these batch sizes) agree with the oracle."""
    wl = synth.workload("traj_dual_arm", scale=0.004)
    oc, pc = oracle_cst(wl), product_cst(wl)
    pc._dev_set_schedule(0, 0, whole_chunks)
    X, off = wl.traj_X, wl.traj_offsets
    vals, jac, cost, grad = pc.traj_cost_grad_batch(X, [int(v) for v in off])
    vref, jref = oc.evaluate_batch(X, 8)
    assert np.abs(vals - vref).max() <= 5e-6
    pd = oc.pair_clearance_batch(X, True, 8)
    clear = (_second_gap(pd) > 1e-5) & (vref[:, 0] < 0.1)
    assert (np.abs(jac[clear, 0] - jref[clear, 0]) <= 1e-4 * np.maximum(np.abs(jref[clear, 0]), 1.0)).all()
    D = X.shape[1]
    for i in range(len(off) - 1):
        cr, gr = _oracle.smooth_cost(X[off[i]:off[i + 1]].ravel(), int(off[i + 1] - off[i]), D, [1.0] * D)
        assert abs(cost[i] - cr) <= 1e-5 * max(1.0, abs(cr))
    v2, _ = pc.evaluate_batch(X[:300])
    assert np.abs(v2 - vref[:300]).max() <= 5e-6


def test_autotuned_schedules_keep_results():
    """Every candidate schedule the autotuner times gives the same edge
    results, and the chosen one is applied to the context."""
    import torch

    from plainmp_amd import autotune
    from plainmp_amd.ompl import BoxMotionValidator

    wl = synth.workload("edges_mixed16", scale=0.125)  # 512 edges
    oc, pc = oracle_cst(wl), product_cst(wl)
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 8)
    mv = BoxMotionValidator(pc, wl.box_width)
    E = len(wl.S1)
    s1 = torch.from_numpy(np.ascontiguousarray(wl.S1.T)).cuda()
    s2 = torch.from_numpy(np.ascontiguousarray(wl.S2.T)).cuda()
    valid = torch.empty(E, dtype=torch.uint8, device="cuda")
    nchk = torch.empty(E, dtype=torch.int32, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    for b in autotune.EDGE_BLOCKS:
        for w in (4, 16, 32):
            autotune.Schedule(b, w).apply(pc)
            mv._dev_check_motions(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
            torch.cuda.synchronize()
            assert np.array_equal(valid.cpu().numpy().astype(bool), ref.astype(bool)), (b, w)
            assert np.array_equal(nchk.cpu().numpy(), n_ref), (b, w)
    sched, rep = autotune.tune_edges(mv, pc, s1.data_ptr(), s2.data_ptr(), E, valid.data_ptr(), nchk.data_ptr(), sh,
                                     repeats=2, cache=None)
    _report("autotune", **rep)
    assert sched.block in autotune.EDGE_BLOCKS and sched.waves_per_cu in autotune.EDGE_WAVES_PER_CU
    got, n_got = mv.check_motions(wl.S1, wl.S2)
    assert np.array_equal(got, ref.astype(bool)) and np.array_equal(n_got, n_ref)


def test_pair_shares_with_a_cloud_and_primitives():
    """A scene of eight primitives and a 4,096-point cloud: the pair
    distances and gradients are the same, bit for bit, under every primitive
    split (the cloud falls in one share), the distances match the oracle's,
    and the adjoint (split over primitive shares too) matches the
    contraction of the full Jacobian tensor."""
    import torch

    from plainmp_amd import autotune
    from plainmp_amd.autograd import pair_clearance

    base = synth.workload("cloud_shelves", scale=128 / 16384)
    wl = synth.Workload("cloud_and_prims", base.robot, synth.mixed_primitives(8, seed=12),
                        cloud=synth.shelf_cloud(4096, seed=13), Q=base.Q)
    oc, pc = oracle_cst(wl), product_cst(wl)
    B, S, P, D = len(wl.Q), pc.n_spheres(), pc.n_prims(), pc.q_dim()
    assert P == 9
    q = torch.from_numpy(np.ascontiguousarray(wl.Q.T)).cuda()
    sh = torch.cuda.current_stream().cuda_stream
    outs = []
    for c in (1,) + autotune.PAIR_SHARES:
        dist = torch.full((S, P, B), float("nan"), dtype=torch.float32, device="cuda")
        grad = torch.full((S, P, D, B), float("nan"), dtype=torch.float32, device="cuda")
        autotune.Schedule(pair_shares=c).apply(pc)
        pc._dev_pair_clearance(q.data_ptr(), B, B, dist.data_ptr(), grad.data_ptr(), sh)
        torch.cuda.synchronize()
        outs.append((c, dist.cpu().numpy(), grad.cpu().numpy()))
    for c, d, g in outs[1:]:
        assert np.array_equal(d, outs[0][1]) and np.array_equal(g, outs[0][2]), c
    dref = oc.pair_clearance_batch(wl.Q, True, 8)
    derr = np.abs(np.moveaxis(outs[0][1], -1, 0) - dref).max()
    assert derr <= 5e-6
    autotune.Schedule().apply(pc)
    qt = torch.tensor(wl.Q, device="cuda", dtype=torch.float64, requires_grad=True)
    d = pair_clearance(pc, qt)
    w = torch.randn(d.shape, device="cuda", dtype=torch.float32, generator=torch.Generator("cuda").manual_seed(1))
    (d * w).sum().backward()
    grad = np.moveaxis(outs[0][2], -1, 0)  # (B, S, P, D)
    ref = np.einsum("bsp,bspj->bj", w.cpu().numpy().astype(np.float64), grad.astype(np.float64))
    scale = np.abs(w.cpu().numpy()).sum(axis=(1, 2))[:, None] * np.abs(grad).max()
    got = qt.grad.cpu().numpy()
    _report("pair_shares_cloud", B=B, P=P, dist_err=derr, vjp_rel=(np.abs(got - ref) / scale).max())
    assert (np.abs(got - ref) <= 1e-5 * scale).all()


def test_pair_shares_keep_results():
    """The pair-distance launch gives the same distances and gradients,
    bit for bit, under every primitive split the autotuner tries, and
    matches the oracle's distances."""
    import torch

    from plainmp_amd import autotune

    wl = synth.workload("grad_mixed32", scale=1 / 16)  # 512 configurations, 32 primitives
    oc, pc = oracle_cst(wl), product_cst(wl)
    B, S, P, D = len(wl.Q), pc.n_spheres(), pc.n_prims(), pc.q_dim()
    q = torch.from_numpy(np.ascontiguousarray(wl.Q.T)).cuda()
    sh = torch.cuda.current_stream().cuda_stream
    outs = []
    for c in (0,) + autotune.PAIR_SHARES:
        dist = torch.full((S, P, B), float("nan"), dtype=torch.float32, device="cuda")
        grad = torch.full((S, P, D, B), float("nan"), dtype=torch.float32, device="cuda")
        autotune.Schedule(pair_shares=c).apply(pc)
        pc._dev_pair_clearance(q.data_ptr(), B, B, dist.data_ptr(), grad.data_ptr(), sh)
        torch.cuda.synchronize()
        outs.append((c, dist.cpu().numpy(), grad.cpu().numpy()))
    for c, d, g in outs[1:]:
        assert np.array_equal(d, outs[0][1]) and np.array_equal(g, outs[0][2]), c
    dref = oc.pair_clearance_batch(wl.Q, True, 8)
    got = np.moveaxis(outs[0][1], -1, 0)
    assert np.abs(got - dref).max() <= 5e-6
    stream = torch.cuda.current_stream()
    sched, rep = autotune.tune_pairs(pc, q.data_ptr(), B, dist.data_ptr(), grad.data_ptr(), stream, repeats=2,
                                     cache=None)
    _report("pair_shares", **{k: v for k, v in rep.items() if k != "all"})
    assert sched.pair_shares in autotune.PAIR_SHARES


def _oracle_link_position(wl, link_name, Q, point=(0.0, 0.0, 0.0)):
    from oracle import urdf_parse

    kin = urdf_parse.kinematic_model(wl.robot.urdf)
    ids = kin.get_joint_ids(list(wl.robot.joint_names))
    lid = kin.get_link_ids([link_name])[0]
    out = np.empty((len(Q), 3))
    for b, q in enumerate(Q):
        kin.set_joint_positions(ids, list(q), True)
        x, y, z, qx, qy, qz, qw = kin.get_link_pose(lid)
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                      [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                      [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
        out[b] = np.array([x, y, z]) + R @ np.asarray(point)
    return out


@pytest.mark.parametrize("point", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.1)])
def test_link_jacobian_vs_central_fd(point):
    """Jacobian of a point fixed to the last link (LinkPoseCst's position
    Jacobian, get_jacobian with rotation ignored) against FP64 central
    differences of the oracle's exact FK."""
    wl = synth.workload("plain_two_bars", scale=1 / 32)
    pc = product_cst(wl)
    link = wl.robot.sphere_specs[-1][0]
    Q = wl.Q
    J = pc.link_jacobian_batch(Q, link, list(point))
    D = Q.shape[1]
    h = 1e-6
    fd = np.empty((len(Q), 3, D))
    for j in range(D):
        e = np.zeros(D)
        e[j] = h
        fd[:, :, j] = (_oracle_link_position(wl, link, Q + e, point) - _oracle_link_position(wl, link, Q - e, point)) / (2 * h)
    err = np.abs(J - fd).max()
    _report("link_jacobian", B=len(Q), point=point, max_err=err)
    assert err <= 1e-5


def test_ik_batched_restarts():
    """Batched multi-start position IK (solve_ik's acceptance test: squared
    position error <= 1e-6 and is_valid) for targets reached by valid
    configurations of the two-bar scene; the answer is checked with the
    oracle's exact FK and validity, and an unreachable target fails."""
    from plainmp_amd.ik import solve_ik_batched

    wl = synth.workload("plain_two_bars", scale=1 / 16)
    oc, pc = oracle_cst(wl), product_cst(wl)
    link = wl.robot.sphere_specs[-1][0]
    ok = oc.is_valid_batch(wl.Q, 8).astype(bool)
    lb, ub = wl.robot.lower, wl.robot.upper
    solved = 0
    for q_star in wl.Q[ok][:4]:
        target = _oracle_link_position(wl, link, q_star[None])[0]
        res = solve_ik_batched(pc, link, target, lb, ub, n_restarts=128, seed=1)
        if not res.success:
            continue
        solved += 1
        err = _oracle_link_position(wl, link, res.q[None])[0] - target
        assert err @ err <= 1e-6
        assert bool(oc.is_valid_batch(res.q[None], 1)[0])
        assert np.all(res.q >= lb - 1e-12) and np.all(res.q <= ub + 1e-12)
        res2 = solve_ik_batched(pc, link, target, lb, ub, n_restarts=128, seed=1)
        assert res2.n_trial == res.n_trial and np.array_equal(res2.q, res.q)
    _report("ik", solved=solved, of=4)
    assert solved >= 3
    far = solve_ik_batched(pc, link, [5.0, 5.0, 5.0], lb, ub, n_restarts=32, seed=0)
    assert not far.success


def _oracle_link_pose(wl, link_name, q):
    from oracle import urdf_parse

    kin = urdf_parse.kinematic_model(wl.robot.urdf)
    kin.set_joint_positions(kin.get_joint_ids(list(wl.robot.joint_names)), list(q), True)
    return np.array(kin.get_link_pose(kin.get_link_ids([link_name])[0]))  # x, y, z, qx, qy, qz, qw


@pytest.mark.parametrize("kind", [6, 7])
def test_ik_batched_pose_targets(kind):
    """Batched IK to position + RPY (6) and position + quaternion (7) targets
    (LinkPoseCst residuals: getRPY() - rpy, quat.coeffs() - q) reached by valid
    configurations of the two-bar scene; answers checked with the oracle's
    exact FK, getRPY and validity (squared residual <= 1e-6)."""
    from plainmp_amd.ik import solve_ik_batched

    wl = synth.workload("plain_two_bars", scale=1 / 16)
    oc, pc = oracle_cst(wl), product_cst(wl)
    link = wl.robot.sphere_specs[-1][0]
    ok = oc.is_valid_batch(wl.Q, 8).astype(bool)
    lb, ub = wl.robot.lower, wl.robot.upper

    def residual(q, target):
        p = _oracle_link_pose(wl, link, q)
        rot = np.array(_oracle.quat_to_rpy(*p[3:])) if kind == 6 else p[3:]
        return np.concatenate([p[:3], rot]) - target

    solved = 0
    for q_star in wl.Q[ok][:4]:
        target = residual(q_star, np.zeros(kind))
        res = solve_ik_batched(pc, link, target, lb, ub, n_restarts=256, seed=2)
        if not res.success:
            continue
        solved += 1
        r = residual(res.q, target)
        assert r @ r <= 1e-6
        assert bool(oc.is_valid_batch(res.q[None], 1)[0])
        assert np.all(res.q >= lb - 1e-12) and np.all(res.q <= ub + 1e-12)
    _report("ik_pose", kind=kind, solved=solved, of=4)
    assert solved >= 3
    with pytest.raises(RuntimeError):
        solve_ik_batched(pc, link, [0.0] * 4, lb, ub, n_restarts=8)
