"""Writes the golden fixtures of this directory (run from the repository root):

    python tests/golden/make_fixtures.py

Each fixture is a seeded, reduced instance of one benchmark workload (robot
and scene come from plainmp_amd.synth by name; the URDF's SHA-256 is stored to
catch generator drift) with the FP64 oracle's outputs.  known_answers.json
holds values the reference's own tests and comments state
(ompl/motion_validator.cpp:23-37, ompl/generate_sequence.py:37-39,
tests/test_psdf.py:89-102,201-204).  cloud_tiles.npz holds a small cloud
and the oracle's tile tables for it.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

from oracle import _oracle  # noqa: E402
from plainmp_amd import synth  # noqa: E402
from support import oracle_cst  # noqa: E402


def urdf_sha(wl):
    return hashlib.sha256(wl.robot.urdf.encode()).hexdigest()


def scene_sha(wl):
    """SHA-256 of the primitive list and the cloud, as the generator produced them."""
    h = hashlib.sha256(json.dumps(wl.prims, sort_keys=True, default=lambda a: np.asarray(a).tolist()).encode())
    if wl.cloud is not None:
        h.update(np.ascontiguousarray(wl.cloud, dtype=np.float64).tobytes())
    return h.hexdigest()


def c1():
    wl = synth.workload("plain_two_bars", scale=1 / 16)  # 64 configurations
    oc = oracle_cst(wl)
    return wl, dict(Q=wl.Q, valid=oc.is_valid_batch(wl.Q, 4).astype(np.uint8),
                    clearance=oc.sphere_clearance_batch(wl.Q, False, 4))


def c2():
    wl = synth.workload("edges_mixed16", scale=1 / 64)  # 64 edges
    oc = oracle_cst(wl)
    ok, n = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, 4)
    return wl, dict(S1=wl.S1, S2=wl.S2, box_width=wl.box_width, valid=ok.astype(np.uint8),
                    n_checked=np.asarray(n, np.int32))


def c3():
    wl = synth.workload("cloud_shelves", scale=16 / 16384)  # 16 configurations
    oc = oracle_cst(wl)
    return wl, dict(Q=wl.Q, valid=oc.is_valid_batch(wl.Q, 4).astype(np.uint8),
                    clearance=oc.sphere_clearance_batch(wl.Q, False, 4))


def c4():
    wl = synth.workload("grad_mixed32", scale=1 / 256)  # 32 configurations
    oc = oracle_cst(wl)
    vals, jac = oc.evaluate_batch(wl.Q, 4)
    return wl, dict(Q=wl.Q, vals=vals, jac=jac, pair=oc.pair_clearance_batch(wl.Q[:8], True, 4))


def c5():
    wl = synth.workload("traj_dual_arm", scale=0.002)  # 4 trajectories
    oc = oracle_cst(wl)
    off, X = wl.traj_offsets, wl.traj_X
    vals, jac = oc.evaluate_batch(X, 4)
    D = X.shape[1]
    w = np.linspace(0.5, 2.0, D)
    cost, grad = [], []
    for i in range(len(off) - 1):
        c, g = _oracle.smooth_cost(X[off[i]:off[i + 1]].ravel(), int(off[i + 1] - off[i]), D, list(w))
        cost.append(c)
        grad.append(np.asarray(g).reshape(-1, D))
    return wl, dict(offsets=off, X=X, weights=w, vals=vals, jac=jac, cost=np.array(cost),
                    grad=np.concatenate(grad))


def cloud_tiles():
    """1,000 points of the shelf cloud generator with a run of duplicates, and
    the oracle's tile tables (field bytes of the 48-byte FP32 records)."""
    P = synth.shelf_cloud(1000, seed=9)
    P[200:260] = P[17]
    t = _oracle.cloud_tiles(P)
    tiles = np.frombuffer(t["tiles"], np.uint8).reshape(-1, 48)[:, :40]
    return dict(points=P, tiles=tiles, **{k: np.frombuffer(t[k], np.uint8) for k in
                                         ("tiles64", "tile_b", "points64", "supers")})


def known_answers():
    return {
        "farthest_sequence_8": [0, 7, 4, 2, 6, 5, 3, 1],
        "farthest_sequence_20": [0, 19, 10, 5, 15, 17, 13, 8, 3, 18, 16, 14, 12, 11, 9, 7, 6, 4, 2, 1],
        "ratio_0.15_travel_rates": [1.0, 0.6, 0.3, 0.9, 0.75, 0.45, 0.15],
        "unit_box_translated_1_2_3": {"lb": [0.5, 1.5, 2.5], "ub": [1.5, 2.5, 3.5]},
        "unit_box_rotate_z_pi_4_lb": [-0.7071067811865476, -0.7071067811865476, -0.5],
        "ground_1_at_origin": 1.0,
        "ground_1_at_x1": 1.0,
    }


def main():
    for name, fn in [("c1_plain_two_bars", c1), ("c2_edges_mixed16", c2), ("c3_cloud_shelves", c3),
                     ("c4_grad_mixed32", c4), ("c5_traj_dual_arm", c5)]:
        wl, arrays = fn()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), workload=np.array(wl.name),
                            urdf_sha256=np.array(urdf_sha(wl)), scene_sha256=np.array(scene_sha(wl)), **arrays)
        print(name, {k: np.shape(v) for k, v in arrays.items()})
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(known_answers(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "cloud_tiles.npz"), **cloud_tiles())


if __name__ == "__main__":
    main()
