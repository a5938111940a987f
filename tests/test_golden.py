"""Golden fixtures (tests/golden, written by tests/golden/make_fixtures.py):
the generator still produces the same robots and scenes, the oracle still
reproduces the stored outputs, and the reference's own known answers hold."""

import json
import os
import sys

import numpy as np
import pytest

from oracle import _oracle
from plainmp_amd import synth
from support import oracle_cst

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_fixtures  # noqa: E402

FIXTURES = ["c1_plain_two_bars", "c2_edges_mixed16", "c3_cloud_shelves", "c4_grad_mixed32", "c5_traj_dual_arm"]
SCALES = {"c1_plain_two_bars": 1 / 16, "c2_edges_mixed16": 1 / 64, "c3_cloud_shelves": 16 / 16384,
          "c4_grad_mixed32": 1 / 256, "c5_traj_dual_arm": 0.002}


def load(name):
    f = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    wl = synth.workload(str(f["workload"]), scale=SCALES[name])
    return f, wl


@pytest.mark.parametrize("name", FIXTURES)
def test_generator_unchanged(name):
    f, wl = load(name)
    assert str(f["urdf_sha256"]) == make_fixtures.urdf_sha(wl)
    assert str(f["scene_sha256"]) == make_fixtures.scene_sha(wl)


def test_oracle_reproduces_c1_c3():
    for name in ("c1_plain_two_bars", "c3_cloud_shelves"):
        f, wl = load(name)
        oc = oracle_cst(wl)
        assert np.array_equal(f["Q"], wl.Q)
        assert np.array_equal(oc.is_valid_batch(f["Q"], 4).astype(np.uint8), f["valid"])
        assert np.allclose(oc.sphere_clearance_batch(f["Q"], False, 4), f["clearance"], rtol=0, atol=1e-12)


def test_oracle_reproduces_c2():
    f, wl = load("c2_edges_mixed16")
    oc = oracle_cst(wl)
    ok, n = oc.check_motions(f["S1"], f["S2"], True, list(f["box_width"]), 0.0, 4)
    assert np.array_equal(ok.astype(np.uint8), f["valid"])
    assert np.array_equal(np.asarray(n, np.int32), f["n_checked"])


def test_oracle_reproduces_c4_c5():
    f, wl = load("c4_grad_mixed32")
    oc = oracle_cst(wl)
    vals, jac = oc.evaluate_batch(f["Q"], 4)
    assert np.allclose(vals, f["vals"], rtol=0, atol=1e-12) and np.allclose(jac, f["jac"], rtol=0, atol=1e-9)
    assert np.allclose(oc.pair_clearance_batch(f["Q"][:8], True, 4), f["pair"], rtol=0, atol=1e-12)
    f, wl = load("c5_traj_dual_arm")
    oc = oracle_cst(wl)
    off, X, w = f["offsets"], f["X"], f["weights"]
    D = X.shape[1]
    for i in range(len(off) - 1):
        c, g = _oracle.smooth_cost(X[off[i]:off[i + 1]].ravel(), int(off[i + 1] - off[i]), D, list(w))
        assert c == pytest.approx(f["cost"][i], rel=1e-12)
        assert np.allclose(g, f["grad"][off[i]:off[i + 1]].ravel(), rtol=1e-12, atol=1e-15)


def test_known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as fh:
        ka = json.load(fh)
    assert _oracle.farthest_sequence(8) == ka["farthest_sequence_8"]
    assert _oracle.farthest_sequence(20) == ka["farthest_sequence_20"]
    assert np.allclose(_oracle.motion_travel_rates(0.15)[1:], ka["ratio_0.15_travel_rates"][1:], atol=1e-12)
    from plainmp_amd import psdf

    b = psdf.BoxSDF([1.0, 1.0, 1.0], psdf.Pose())
    b.translate(np.array([1.0, 2.0, 3.0]))
    assert np.allclose(b.lb, ka["unit_box_translated_1_2_3"]["lb"]) and np.allclose(b.ub, ka["unit_box_translated_1_2_3"]["ub"])
    b = psdf.BoxSDF([1.0, 1.0, 1.0], psdf.Pose())
    b.rotate_z(np.pi / 4)
    assert np.allclose(b.lb, ka["unit_box_rotate_z_pi_4_lb"])
    g = _oracle.Prim.ground(1.0)  # the product's evaluate runs on the GPU: tests/test_gpu_golden.py
    assert g.evaluate(np.zeros(3)) == ka["ground_1_at_origin"]
    assert g.evaluate(np.array([1.0, 0.0, 0.0])) == ka["ground_1_at_x1"]


def test_oracle_reproduces_cloud_tiles():
    """The oracle's tile tables of the stored cloud equal the stored ones."""
    f = np.load(os.path.join(GOLDEN, "cloud_tiles.npz"), allow_pickle=False)
    t = _oracle.cloud_tiles(f["points"])
    assert np.array_equal(np.frombuffer(t["tiles"], np.uint8).reshape(-1, 48)[:, :40], f["tiles"])
    for k in ("tiles64", "tile_b", "points64", "supers"):
        assert np.array_equal(np.frombuffer(t[k], np.uint8), f[k]), k
