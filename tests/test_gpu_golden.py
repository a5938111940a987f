"""The HIP path against the committed golden fixtures (tests/golden): the
stored oracle outputs of reduced C1-C5 instances and the reference's known
answers.  Bars as in test_gpu_parity: booleans and counts exact, FP32 values
within 5e-6 m, Jacobian entries within 1e-4 relative where the argmin is
unambiguous, smoothness cost / gradient within 1e-5 relative."""

import json
import os

import numpy as np
import pytest

from plainmp_amd import synth
from support import product_cst

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCALES = {"c1_plain_two_bars": 1 / 16, "c2_edges_mixed16": 1 / 64, "c3_cloud_shelves": 16 / 16384,
          "c4_grad_mixed32": 1 / 256, "c5_traj_dual_arm": 0.002}


def load(name):
    f = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return f, synth.workload(str(f["workload"]), scale=SCALES[name])


@pytest.mark.parametrize("name", ["c1_plain_two_bars", "c3_cloud_shelves"])
def test_golden_validity_and_clearance(name):
    f, wl = load(name)
    pc = product_cst(wl)
    assert np.array_equal(pc.is_valid_batch(f["Q"]).astype(np.uint8), f["valid"])
    assert np.abs(pc.sphere_clearance_batch(f["Q"], False) - f["clearance"]).max() <= 5e-6


def test_golden_edges():
    from plainmp_amd.ompl import BoxMotionValidator

    f, wl = load("c2_edges_mixed16")
    ok, n = BoxMotionValidator(product_cst(wl), f["box_width"]).check_motions(f["S1"], f["S2"])
    assert np.array_equal(ok.astype(np.uint8), f["valid"])
    assert np.array_equal(n, f["n_checked"])


def test_golden_evaluate_and_pairs():
    f, wl = load("c4_grad_mixed32")
    pc = product_cst(wl)
    vals, jac = pc.evaluate_batch(f["Q"])
    assert np.abs(vals - f["vals"]).max() <= 5e-6
    s = np.sort(f["pair"].reshape(8, -1), axis=1)
    clear = np.zeros(len(f["Q"]), bool)
    clear[:8] = s[:, 1] - s[:, 0] > 1e-5
    assert (np.abs(jac - f["jac"])[clear] <= 1e-4 * np.maximum(np.abs(f["jac"][clear]), 1.0)).all()
    dist, _ = pc.pair_clearance_batch(f["Q"][:8], False)
    assert np.abs(dist - f["pair"]).max() <= 5e-6


def test_golden_trajectories():
    f, wl = load("c5_traj_dual_arm")
    pc = product_cst(wl)
    vals, jac, cost, grad = pc.traj_cost_grad_batch(f["X"], [int(v) for v in f["offsets"]], list(f["weights"]))
    assert np.abs(vals - f["vals"]).max() <= 5e-6
    assert np.allclose(cost, f["cost"], rtol=1e-5, atol=1e-6)
    assert np.allclose(grad, f["grad"], rtol=1e-5, atol=1e-6)


def test_golden_known_answers_on_gpu():
    from plainmp_amd import psdf

    with open(os.path.join(GOLDEN, "known_answers.json")) as fh:
        ka = json.load(fh)
    g = psdf.GroundSDF(1.0)
    assert g.evaluate(np.zeros(3)) == ka["ground_1_at_origin"]
    assert g.evaluate(np.array([1.0, 0.0, 0.0])) == ka["ground_1_at_x1"]


def test_golden_cloud_tiles():
    """The device build of the stored cloud's tile tables equals the stored
    tables byte for byte."""
    from plainmp_amd import kdtree

    f = np.load(os.path.join(GOLDEN, "cloud_tiles.npz"), allow_pickle=False)
    t = kdtree.build_cloud_tiles(f["points"])
    assert np.array_equal(np.frombuffer(t["tiles"], np.uint8).reshape(-1, 48)[:, :40], f["tiles"])
    for k in ("tiles64", "tile_b", "points64", "supers"):
        assert np.array_equal(np.frombuffer(t[k], np.uint8), f[k]), k
