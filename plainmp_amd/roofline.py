"""Algorithmic work model of the FK + sphere-SDF path, for roofline figures.

Convention: every add, sub, mul, abs, min, max, compare and sqrt counts 1
flop and an FMA counts 2; work is full evaluation (no early exit, no
culling).  Per-block costs:

    revolute joint (sincos 25 + local quat 3 + origin product 28 + compose 61)   117
        (89 when the origin rotation is flagged identity)
    prismatic joint                                                               39
    fixed link compose                                                            61 (33 identity)
    rotation matrix per sphere-bearing link                                        24
    sphere world centre R p + t                                                    18
    sphere-primitive pair: box aligned / box rotated / sphere / cylinder z /
        cylinder rotated / ground                                                 25 / 40 / 11 / 23 / 38 / 4
    self sphere pair                                                               11
    adjoint per (sphere, primitive) pair: analytic gradient 12 + 17 per column
    cloud pair as the K = 4 product [x, 1] . [-2y, |y|^2]                           8
    smoothness per waypoint per dof (value + gradient)                             16
    lerp per waypoint per dof                                                       3

MI355X peaks (AMD published): FP32 vector 157.3 TFLOP/s (78.6 without packed
FP32), FP64 vector 78.6, FP32 matrix 157.3, HBM3E 8 TB/s.
"""

from __future__ import annotations

import numpy as np

PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP32_VALU_NONPACKED_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
N_SIMD = 1024  # 256 CUs x 4 SIMDs
CLOCK_GHZ = 2.4  # peak engine clock
MFMA_F32_16X16X4_FLOP = 2 * 16 * 16 * 4  # one v_mfma_f32_16x16x4f32


def _pair_cost(p) -> int:
    t = p["type"]
    if t == "box":
        return 25 if np.allclose(p["rotation"], np.eye(3)) else 40
    if t == "sphere":
        return 11
    if t == "cylinder":
        R = np.asarray(p["rotation"])
        z_aligned = abs(R[2, 2] - 1) < 1e-6 and abs(R[0, 2]) < 1e-6 and abs(R[1, 2]) < 1e-6
        return 23 if z_aligned else 38
    if t == "ground":
        return 4
    raise KeyError(t)


def fk_flops(robot) -> int:
    """Joint + fixed-link compose + rotation matrices + sphere centres of one configuration."""
    import xml.etree.ElementTree as ET

    root = ET.fromstring(robot.urdf)
    total = 0
    for j in root.findall("joint"):
        jt = j.get("type")
        o = j.find("origin")
        rpy = [float(v) for v in o.get("rpy", "0 0 0").split()] if o is not None else [0, 0, 0]
        identity = all(abs(v) < 2e-6 for v in rpy)
        if jt in ("revolute", "continuous"):
            total += 89 if identity else 117
        elif jt == "prismatic":
            total += 39
        else:
            total += 33 if identity else 61
    links = {s[0] for s in robot.sphere_specs}
    total += 24 * len(links) + 18 * robot.n_spheres
    return total


def config_flops(robot, prims, self_pairs: int = 0) -> int:
    """Full-evaluation flops of one collision check (FK + all pairs)."""
    pair = sum(_pair_cost(p) for p in prims)
    return fk_flops(robot) + robot.n_spheres * pair + 11 * self_pairs


def waypoint_flops(robot, prims, self_pairs: int = 0) -> int:
    """One edge waypoint: lerp of every joint plus a collision check."""
    return 3 * robot.dof + config_flops(robot, prims, self_pairs)


def evaluate_flops(robot, prims, self_pairs: int = 0) -> int:
    """Reference-semantics evaluate of one configuration (constraint values
    and rows, primitive_sphere_collision.cpp:256-399): every pair's value plus
    the Jacobian row of the external and of the self argmin (analytic SDF
    gradient 12 + 17 per column each)."""
    rows = (1 if prims else 0) + (1 if self_pairs else 0)
    return config_flops(robot, prims, self_pairs) + rows * (12 + 17 * robot.dof)


def traj_waypoint_flops(robot, prims, self_pairs: int = 0) -> int:
    """One trajectory waypoint of the cost / gradient launch: evaluate plus
    the smoothness value and gradient (16 per dof)."""
    return evaluate_flops(robot, prims, self_pairs) + 16 * robot.dof


def sources_digest() -> str:
    """SHA-256 of the kernel and C-ABI sources, to tie measured figures (the
    committed traffic table) to the build they were measured on."""
    import hashlib
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(here)):
        if f.endswith((".hip", ".hpp", ".h", ".cpp")):
            with open(os.path.join(here, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def pair_grad_flops(robot, prims, n_rel_mean: float) -> float:
    """Per configuration: FK + every (sphere, primitive) value and its adjoint."""
    n_pairs = robot.n_spheres * len(prims)
    return fk_flops(robot) + sum(robot.n_spheres * _pair_cost(p) for p in prims) + n_pairs * (12 + 17 * n_rel_mean)


# ------------------------------------------------- HBM traffic from counters
def _counter_rows(path):
    """Rows of a rocprofv3 counter-collection CSV (a file, or a directory
    searched for *counter_collection.csv)."""
    import csv
    import glob
    import os

    if os.path.isdir(path):
        hits = sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
        if not hits:
            raise FileNotFoundError(f"no counter_collection.csv under {path}")
        path = hits[0]
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _kernel_base(name: str) -> str:
    """A kernel's name without return type, namespaces and argument list
    (template arguments kept): 'void pmp::edges_kernel<false>(...)' ->
    'edges_kernel<false>'."""
    n = name.replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth, cut, last_sep = 0, len(n), 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
        elif ch == ":" and depth == 0 and n[i + 1:i + 2] == ":":
            last_sep = i + 2
    return n[last_sep:cut]


def _trace_rows(path):
    """Rows of a rocprofv3 kernel-trace CSV (a file, or a directory searched
    for *kernel_trace.csv)."""
    import csv
    import glob
    import os

    if os.path.isdir(path):
        hits = sorted(glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True))
        if not hits:
            raise FileNotFoundError(f"no kernel_trace.csv under {path}")
        path = hits[0]
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def matrix_core_table(mfma_path, trace_path) -> dict:
    """Per kernel and grid (threads): the FP32 MFMA instructions and matrix-
    core busy cycles per dispatch (a `--pmc SQ_INSTS_VALU_MFMA_F32
    SQ_VALU_MFMA_BUSY_CYCLES` run) and the average dispatch duration in µs
    (a `--kernel-trace` run), for kernels that issue MFMA."""
    import statistics

    acc = {}
    for r in _counter_rows(mfma_path):
        name = r.get("Counter_Name")
        if name not in ("SQ_INSTS_VALU_MFMA_F32", "SQ_VALU_MFMA_BUSY_CYCLES"):
            continue
        key = (_kernel_base(r["Kernel_Name"]), int(r["Grid_Size"]))
        acc.setdefault(key, {}).setdefault(name, []).append(float(r["Counter_Value"]))
    dur = {}
    for r in _trace_rows(trace_path):
        g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        key = (_kernel_base(r["Kernel_Name"]), g)
        dur.setdefault(key, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    out = {}
    for (k, g), cs in sorted(acc.items()):
        insts = statistics.mean(cs.get("SQ_INSTS_VALU_MFMA_F32", [0.0]))
        if insts <= 0 or (k, g) not in dur:
            continue
        out.setdefault(k, {})[str(g)] = {
            "mfma_f32_insts": round(insts, 1),
            "mfma_busy_cycles": round(statistics.mean(cs.get("SQ_VALU_MFMA_BUSY_CYCLES", [0.0])), 1),
            "duration_us": round(statistics.mean(dur[(k, g)]), 3), "launches": len(dur[(k, g)])}
    return out


def traffic_table(fetch_path, write_path, calib_mb: int) -> dict:
    """Per kernel and grid, HBM bytes per launch from a FETCH_SIZE run and a
    WRITE_SIZE run (KiB per dispatch), each scaled by what the same counter
    reads for the plain device copy of `calib_mb` MiB the run made
    (bench.py --calib-copy-mb): the copy moves calib_mb MiB each way."""
    import statistics

    def per_kernel(rows, counter):
        acc = {}
        for r in rows:
            if r.get("Counter_Name") != counter:
                continue
            key = (_kernel_base(r["Kernel_Name"]), int(r["Grid_Size"]))
            acc.setdefault(key, []).append(float(r["Counter_Value"]))
        return acc

    fetch = per_kernel(_counter_rows(fetch_path), "FETCH_SIZE")
    write = per_kernel(_counter_rows(write_path), "WRITE_SIZE")
    kib = calib_mb * 1024.0

    def calib(acc):  # the launch whose counter reads closest to the copy size (the copy itself)
        best = None
        for (k, g), v in acc.items():
            m = max(v)
            if 0.25 * kib <= m <= 2.0 * kib and (best is None or abs(m - kib) < abs(best - kib)):
                best = m
        if best is None:
            raise ValueError("calibration copy not found in the counter run")
        return kib / best

    ff, wf = calib(fetch), calib(write)
    out = {"_calibration": {"copy_MiB": calib_mb, "fetch_factor": ff, "write_factor": wf,
                            "sources_digest": sources_digest()}}
    for (k, g) in sorted(set(fetch) & set(write)):
        fv, wv = statistics.mean(fetch[(k, g)]), statistics.mean(write[(k, g)])
        rec = out.setdefault(k, {"grids": {}})
        rec["grids"][str(g)] = {"launches": len(fetch[(k, g)]), "fetch_KiB_raw": round(fv, 2),
                                "write_KiB_raw": round(wv, 2),
                                "bytes_per_launch": int(round((fv * ff + wv * wf) * 1024))}
    for k, rec in out.items():
        if k.startswith("_"):
            continue
        grids = rec["grids"]
        big = max(grids, key=lambda g: int(g))  # the largest grid of a kernel is its benchmark launch
        rec["bytes_per_launch"] = grids[big]["bytes_per_launch"]
    return out


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="HBM traffic per kernel launch from two rocprofv3 --pmc runs")
    ap.add_argument("--fetch", required=True, help="output of the FETCH_SIZE run (CSV or directory)")
    ap.add_argument("--write", required=True, help="output of the WRITE_SIZE run (CSV or directory)")
    ap.add_argument("--calib-mb", type=int, default=1024)
    ap.add_argument("--mfma", help="output of the MFMA counter run (optional, with --trace)")
    ap.add_argument("--trace", help="output of the kernel-trace run (optional, with --mfma)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    table = traffic_table(a.fetch, a.write, a.calib_mb)
    if a.mfma and a.trace:
        table["_matrix_cores"] = matrix_core_table(a.mfma, a.trace)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
