"""Benchmark of the MI355X collision path on the batched edge-validation workload.

Workload (one step): 4,096 edges between uniform configurations of a
Panda-class 7R arm with 60 spheres, among 16 mixed primitives, box motion
validator at 0.05 rad (1-117 waypoints per edge), validated by one HIP-graph
replay of the persistent edge kernel.  Inputs are resident in HBM.

`value` counts collision-checked configurations per second as the reference
counts them: the states checkMotion evaluates in its visiting order, up to
and including each edge's first invalid state (the kernel reports this count
per edge, exact).  The same line carries edges/s, the other workloads' rates,
the roofline figures of the edge kernel (its duration from HIP events
recorded around it on the stream, in direct launches of the same work) and the CPU baseline (the FP64 oracle on
this box's host cores, on a bounded sample, counted the same way).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def _soa(t, x):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x.T)).to("cuda")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--calib-copy-mb", type=int, default=0, help="run one device copy of this size (counter calibration)")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file (default: none, i.e. tune in this run)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every configuration, edge and trajectory is independent, so several GPUs "
              "would each run this same kernel on a contiguous slice; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import plainmp_amd  # noqa: F401
    from plainmp_amd import autotune, roofline, synth
    from plainmp_amd.ompl import BoxMotionValidator
    from plainmp_amd.synth import build_constraint as product_cst

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    # ------------------------------------------------ headline: batched edges
    wl = synth.workload("edges_mixed16")
    cst = product_cst(wl)
    mv = BoxMotionValidator(cst, wl.box_width)
    E = len(wl.S1)
    with torch.cuda.stream(stream):
        s1 = _soa(torch, wl.S1)
        s2 = _soa(torch, wl.S2)
        valid = torch.empty(E, dtype=torch.uint8, device=dev)
        nchk = torch.empty(E, dtype=torch.int32, device=dev)
        if args.calib_copy_mb:  # a plain device copy of known size, for scaling HBM byte counters
            a = torch.empty(args.calib_copy_mb * (1 << 20) // 4, dtype=torch.float32, device=dev)
            a.fill_(1.0)
            b = torch.empty_like(a)
            b.copy_(a)
            del a, b
        # schedule: the autotuner's choice for this device and workload (timed outside the measured region)
        schedule, tune_report = autotune.tune_edges(mv, cst, s1.data_ptr(), s2.data_ptr(), E, valid.data_ptr(),
                                                    nchk.data_ptr(), sh, key="edges_mixed16", cache=args.tune_cache)
        graph = mv._dev_capture_graph(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
        for _ in range(args.warmup):
            graph.launch()
    stream.synchronize()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(args.steps):
            graph.launch()
        ev1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    # the edge kernel alone: the same launch made directly on the stream, with
    # HIP events recorded around the kernel itself
    kernel_ms = []
    for _ in range(max(args.steps, 5)):
        mv._dev_check_motions(s1.data_ptr(), s2.data_ptr(), E, E, valid.data_ptr(), nchk.data_ptr(), sh)
        kernel_ms.append(mv._dev_edge_kernel_ms(sh))
    kernel_ms = float(np.mean(kernel_ms))
    stream.synchronize()
    n_recheck = mv.last_recheck_count()
    # the graph's own outputs of its last replay, checked against the oracle
    # after the timed region (direct calls above rewrite them with the same work)
    graph.launch()
    stream.synchronize()
    checked = int(nchk.sum().item())
    ok = valid.cpu().numpy()
    parity = _parity_headline(wl, ok, nchk.cpu().numpy())
    edges_per_s = E / (ms_step * 1e-3)
    configs_per_s = checked / (ms_step * 1e-3)
    n_wp_total = _total_waypoints(wl)
    wp_flops = roofline.waypoint_flops(wl.robot, wl.prims)
    achieved_tflops = checked * wp_flops / (kernel_ms * 1e-3) / 1e12
    traffic, traffic_note = _traffic_per_launch("edges_kernel<false>")

    out = {
        "metric": "collision-checked configs/sec + Jacobian evals/sec on 1 MI355X vs plain GPU & CPU",
        "value": configs_per_s,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "kernel_ms": kernel_ms,
        "wall_s_timed": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 in-band re-decision, f64 edge step ratio and waypoint lerp)",
        "data": "synthetic (seeded robot, scene and edges; see plainmp_amd/synth.py)",
        "config": {
            "workload": "edges_mixed16: Panda-class 7R, 60 spheres, 16 mixed primitives, 4096 edges, "
                        "box validator 0.05 rad, one HIP-graph replay per step",
            "edges": E,
            "waypoints_total": n_wp_total,
            "waypoints_checked": checked,
            "edges_valid_frac": float(ok.astype(bool).mean()),
            "edges_rechecked_fp64": n_recheck,
        },
        "parity_checked": parity["equal"],
        "parity": parity,
        "edges_per_s": edges_per_s,
        "schedule": {**schedule.as_dict(), "autotune": tune_report},
        "covered_configs_per_s": n_wp_total / (ms_step * 1e-3),
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": roofline.PEAK_FP32_VALU_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / roofline.PEAK_FP32_VALU_TFLOPS,
            "traffic": traffic,
            "traffic_note": traffic_note,
            "flops_per_waypoint": wp_flops,
            "note": "algorithmic flops = reference-order checked states x full-evaluation flops per state "
                    "(no culling, no early exit inside a state) / edge-kernel time; FP32 VALU peak with "
                    "packed FP32, which this build leaves off (78.6 TFLOP/s without it)",
        },
    }

    if not args.no_extra:
        out["extra"] = _extra(torch, stream, dev)
        out["rooflines"] = out["extra"].pop("_rooflines")
        out["plain_per_call_gpu"] = _plain_per_call(wl, mv)
        out["speedup_vs_plain_per_call_edges"] = edges_per_s / out["plain_per_call_gpu"]["edges_per_s"]
        # north-star comparison: Panda-class arm vs 32 primitives, batched
        # configurations against one configuration per call (BASELINE "B")
        b32 = _plain_per_call_configs()
        out["plain_per_call_gpu_configs"] = b32
        out["speedup_vs_plain_per_call_configs"] = (out["extra"]["grad_mixed32_is_valid"]["configs_per_s"]
                                                    / b32["configs_per_s"])
    if not args.no_cpu:
        out["cpu_baseline"] = _cpu_baseline(wl)
        out["speedup_vs_cpu_configs"] = configs_per_s / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_one_thread_configs"] = configs_per_s / out["cpu_baseline"]["one_thread"]["value"]
        if not args.no_extra:
            out["cpu_baseline_evaluate"] = _cpu_evaluate_baseline()
            out["speedup_vs_cpu_jacobians"] = (out["extra"]["grad_mixed32_evaluate"]["jacobians_per_s"]
                                               / out["cpu_baseline_evaluate"]["value"])
    print(json.dumps(out), flush=True)


def _traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC table
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by a plain copy's reading),
    or None when the table was measured on other kernel sources than these."""
    from plainmp_amd import roofline

    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, "no traffic table"
    with open(path) as f:
        table = json.load(f)
    if table.get("_calibration", {}).get("sources_digest") != roofline.sources_digest():
        return None, "traffic table measured on other kernel sources (stale): not used"
    rec = table.get(kernel)
    if rec is None:
        return None, "kernel not in the traffic table"
    return rec.get("bytes_per_launch"), "profiles/traffic.json, this build's sources"


def _matrix_cores(kernel):
    """Matrix-core counters and rocprof duration of `kernel`'s largest grid
    from the committed table, or None when it was measured on other kernel
    sources than these."""
    from plainmp_amd import roofline

    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    if table.get("_calibration", {}).get("sources_digest") != roofline.sources_digest():
        return None
    grids = table.get("_matrix_cores", {}).get(kernel)
    if not grids:
        return None
    return grids[max(grids, key=int)]


def _parity_headline(wl, got_valid, got_n):
    """The headline replay's per-edge validity and reference-order checked
    counts against the FP64 oracle (checkMotion, motion_validator.cpp:17-79)
    on every host thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from support import oracle_cst

    oc = oracle_cst(wl)
    t0 = time.perf_counter()
    ref, n_ref = oc.check_motions(wl.S1, wl.S2, True, list(wl.box_width), 0.0, _host_threads())
    v_eq = bool(np.array_equal(got_valid.astype(bool), ref.astype(bool)))
    n_eq = bool(np.array_equal(got_n, n_ref))
    if not (v_eq and n_eq):
        print(f"PARITY FAILURE: valid equal {v_eq}, n_checked equal {n_eq}", file=sys.stderr, flush=True)
    return {"equal": v_eq and n_eq, "valid_equal": v_eq, "n_checked_equal": n_eq, "edges": int(len(ref)),
            "oracle_s": time.perf_counter() - t0}


def _host_threads():
    return max(1, len(os.sched_getaffinity(0)))


def _host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def _plain_per_call_configs(n=400):
    """One configuration per call on the Panda-class arm against 32 primitives
    (grad_mixed32): synchronous host->device copy, the validity launches and
    the device->host copy, as a per-call drop-in would do."""
    from plainmp_amd import synth
    from plainmp_amd.synth import build_constraint as product_cst

    wl = synth.workload("grad_mixed32")
    cst = product_cst(wl)
    for q in wl.Q[:20]:
        cst.is_valid(q)
    t0 = time.perf_counter()
    for q in wl.Q[:n]:
        cst.is_valid(q)
    dt = time.perf_counter() - t0
    return {"configs": n, "s": dt, "configs_per_s": n / dt, "workload": "grad_mixed32 (7R, 60 spheres, 32 primitives)"}


def _total_waypoints(wl):
    from plainmp_amd.synth import edge_waypoint_counts

    return int(edge_waypoint_counts(wl.S1, wl.S2, wl.box_width).sum())


def _time_stream(torch, stream, fn, steps=20, warmup=3):
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            fn()
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / steps


def _extra(torch, stream, dev):
    from plainmp_amd import autotune, roofline, synth
    from plainmp_amd.synth import build_constraint as product_cst

    res = {}
    sh = stream.cuda_stream
    # plain_two_bars: 1,024 configs, is_valid
    wl = synth.workload("plain_two_bars")
    cst = product_cst(wl)
    B = len(wl.Q)
    with torch.cuda.stream(stream):
        q = _soa(torch, wl.Q)
        valid = torch.empty(B, dtype=torch.uint8, device=dev)
    ms = _time_stream(torch, stream, lambda: cst._dev_is_valid(q.data_ptr(), B, B, valid.data_ptr(), 0, sh))
    res["plain_two_bars_is_valid"] = {"configs": B, "ms": ms, "configs_per_s": B / (ms * 1e-3),
                                      "valid_frac": float(valid.float().mean().item())}
    # grad_mixed32: 8,192 configs, every (sphere, primitive) distance + d/dq
    wl = synth.workload("grad_mixed32")
    cst = product_cst(wl)
    B, S, P, D = len(wl.Q), cst.n_spheres(), cst.n_prims(), cst.q_dim()
    with torch.cuda.stream(stream):
        q = _soa(torch, wl.Q)
        dist = torch.empty((S, P, B), dtype=torch.float32, device=dev)
        grad = torch.empty((S, P, D, B), dtype=torch.float32, device=dev)
    _, tune_pair = autotune.tune_pairs(cst, q.data_ptr(), B, dist.data_ptr(), grad.data_ptr(), stream,
                                       key="grad_mixed32", cache=None)
    ms = _time_stream(torch, stream,
                      lambda: cst._dev_pair_clearance(q.data_ptr(), B, B, dist.data_ptr(), grad.data_ptr(), sh))
    bytes_out = (dist.numel() + grad.numel()) * 4 + q.numel() * 8
    res["grad_mixed32_pair_jacobian"] = {
        "configs": B, "ms": ms, "jacobians_per_s": B / (ms * 1e-3),
        "hbm_gbs_algorithmic": bytes_out / (ms * 1e-3) / 1e9,
        "frac_of_hbm_peak": bytes_out / (ms * 1e-3) / 1e9 / roofline.PEAK_HBM_GBS,
        "autotune": tune_pair,
    }
    rl = {}
    rl["pair_kernel"] = {"workload": "grad_mixed32 pair distances + d/dq, 8192 configs", "bound": "hbm",
                         "achieved": bytes_out / (ms * 1e-3) / 1e9, "peak": roofline.PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": bytes_out / (ms * 1e-3) / 1e9 / roofline.PEAK_HBM_GBS,
                         "bytes_per_launch": bytes_out, "launch_ms": ms,
                         "note": "algorithmic bytes: dist S x P x B + grad S x P x D x B (FP32) written, q read"}
    # validity of the same 8,192 configurations (the north-star batched rate,
    # against one configuration per call below)
    with torch.cuda.stream(stream):
        valid = torch.empty(B, dtype=torch.uint8, device=dev)
    ms_v = _time_stream(torch, stream, lambda: cst._dev_is_valid(q.data_ptr(), B, B, valid.data_ptr(), 0, sh))
    res["grad_mixed32_is_valid"] = {"configs": B, "ms": ms_v, "configs_per_s": B / (ms_v * 1e-3),
                                    "valid_frac": float(valid.float().mean().item())}
    # reference-semantics evaluate (value + 1 Jacobian row) on the same batch
    C = cst.cst_dim()
    with torch.cuda.stream(stream):
        vals = torch.empty((C, B), dtype=torch.float32, device=dev)
        jac = torch.empty((C, D, B), dtype=torch.float32, device=dev)
    _, tune_eval = autotune.tune_evaluate(cst, q.data_ptr(), B, vals.data_ptr(), jac.data_ptr(), stream,
                                          key="grad_mixed32", cache=None)
    ms = _time_stream(torch, stream, lambda: cst._dev_evaluate(q.data_ptr(), B, B, vals.data_ptr(), jac.data_ptr(), sh))
    res["grad_mixed32_evaluate"] = {"configs": B, "ms": ms, "jacobians_per_s": B / (ms * 1e-3),
                                    "autotune": tune_eval}
    ev_flops = roofline.evaluate_flops(wl.robot, wl.prims)
    rl["eval_group_kernel"] = {"workload": "grad_mixed32 evaluate (value + row), 8192 configs", "bound": "valu",
                               "achieved": B * ev_flops / (ms * 1e-3) / 1e12, "peak": roofline.PEAK_FP32_VALU_TFLOPS,
                               "unit": "TFLOP/s", "frac": B * ev_flops / (ms * 1e-3) / 1e12 / roofline.PEAK_FP32_VALU_TFLOPS,
                               "flops_per_config": ev_flops, "launch_ms": ms}
    # adjoint (VJP) used by the autograd backward
    with torch.cuda.stream(stream):
        gd = torch.ones((S, P, B), dtype=torch.float32, device=dev)
        gq = torch.empty((D, B), dtype=torch.float32, device=dev)
    ms = _time_stream(torch, stream,
                      lambda: cst._dev_pair_clearance_vjp(q.data_ptr(), B, B, gd.data_ptr(), gq.data_ptr(), sh))
    res["grad_mixed32_vjp"] = {"configs": B, "ms": ms, "vjps_per_s": B / (ms * 1e-3)}
    # traj_dual_arm: 2,048 trajectories of 20-100 waypoints, one launch: evaluate rows
    # of every waypoint + smoothness cost and gradient
    wl = synth.workload("traj_dual_arm")
    cst = product_cst(wl)
    off, X = wl.traj_offsets, wl.traj_X
    N, D, C, T = len(X), cst.q_dim(), cst.cst_dim(), len(off) - 1
    offl = [int(v) for v in off]
    with torch.cuda.stream(stream):
        x = _soa(torch, X)
        vals = torch.empty((C, N), dtype=torch.float32, device=dev)
        jac = torch.empty((C, D, N), dtype=torch.float32, device=dev)
        cost = torch.empty(T, dtype=torch.float32, device=dev)
        grad = torch.empty((D, N), dtype=torch.float32, device=dev)
    ms = _time_stream(torch, stream, lambda: cst._dev_traj_cost_grad(
        x.data_ptr(), N, N, offl, [], vals.data_ptr(), jac.data_ptr(), cost.data_ptr(), grad.data_ptr(), sh))
    res["traj_dual_arm_cost_grad"] = {"trajectories": T, "waypoints": N, "ms": ms,
                                      "waypoints_per_s": N / (ms * 1e-3), "trajectories_per_s": T / (ms * 1e-3)}
    tw = roofline.traj_waypoint_flops(wl.robot, wl.prims, synth.self_pair_count(wl.robot))
    rl["eval_units_kernel"] = {"workload": "traj_dual_arm cost + gradient, 122138 waypoints", "bound": "valu",
                               "achieved": N * tw / (ms * 1e-3) / 1e12, "peak": roofline.PEAK_FP32_VALU_TFLOPS,
                               "unit": "TFLOP/s", "frac": N * tw / (ms * 1e-3) / 1e12 / roofline.PEAK_FP32_VALU_TFLOPS,
                               "flops_per_waypoint": tw, "launch_ms": ms}
    # the same trajectories as the SQP solver's sparse inequality sequence:
    # evaluate rows of every waypoint + motion-step box rows, CSC per trajectory
    msbox = [0.1] * D
    r_off, z_off, p_off = cst.traj_sqp_sizes(offl, True)
    with torch.cuda.stream(stream):
        cv = torch.empty(max(r_off[-1], 1), dtype=torch.float64, device=dev)
        cp = torch.empty(max(p_off[-1], 1), dtype=torch.int32, device=dev)
        ri = torch.empty(max(z_off[-1], 1), dtype=torch.int32, device=dev)
        va = torch.empty(max(z_off[-1], 1), dtype=torch.float64, device=dev)
    ms = _time_stream(torch, stream, lambda: cst._dev_traj_sqp_ineq(
        x.data_ptr(), N, N, offl, msbox, cv.data_ptr(), cp.data_ptr(), ri.data_ptr(), va.data_ptr(), sh))
    out_bytes = r_off[-1] * 8 + p_off[-1] * 4 + z_off[-1] * 12
    res["traj_dual_arm_sqp_csc"] = {"trajectories": T, "waypoints": N, "rows": r_off[-1], "nnz": z_off[-1],
                                    "ms": ms, "nnz_per_s": z_off[-1] / (ms * 1e-3),
                                    "output_gb_per_s": out_bytes / (ms * 1e-3) / 1e9,
                                    "note": "includes the evaluate rows of every waypoint (same launch as "
                                            "traj_dual_arm_cost_grad without the smoothness terms)"}
    # cloud_shelves: 16,384 configurations of an 8-DOF arm with 90 spheres against a
    # 65,536-point cloud (sphere-vs-cloud distances on the matrix cores)
    wl = synth.workload("cloud_shelves")
    cst = product_cst(wl)
    B = len(wl.Q)
    with torch.cuda.stream(stream):
        q = _soa(torch, wl.Q)
        valid = torch.empty(B, dtype=torch.uint8, device=dev)
    ms = _time_stream(torch, stream, lambda: cst._dev_is_valid(q.data_ptr(), B, B, valid.data_ptr(), 0, sh), steps=10)
    n_pts = len(wl.cloud)
    res["cloud_shelves_is_valid"] = {
        "configs": B, "cloud_points": n_pts, "spheres": cst.n_spheres(), "ms": ms, "configs_per_s": B / (ms * 1e-3),
        "valid_frac": float(valid.float().mean().item()),
        "note": "whole validity call (FP32 walk, matrix-core cloud pass, decisions, FP64 re-decisions); the "
                "cloud kernel's matrix-core utilisation is measured with counters (profiles/)",
    }
    mc = _matrix_cores("cloud_min_kernel<true>")
    if mc is not None:
        flop = mc["mfma_f32_insts"] * roofline.MFMA_F32_16X16X4_FLOP
        simd_cycles = roofline.N_SIMD * mc["duration_us"] * 1e-6 * roofline.CLOCK_GHZ * 1e9
        rl["cloud_min_kernel"] = {
            "workload": "cloud_shelves: 16,384 configs x 90 spheres vs 65,536 points", "bound": "mfma",
            "achieved": flop / (mc["duration_us"] * 1e-6) / 1e12, "peak": roofline.PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s", "frac": flop / (mc["duration_us"] * 1e-6) / 1e12 / roofline.PEAK_FP32_MFMA_TFLOPS,
            "mfma_busy": mc["mfma_busy_cycles"] / simd_cycles, "mfma_insts": mc["mfma_f32_insts"],
            "launch_us": mc["duration_us"],
            "note": "executed matrix-core work (SQ_INSTS_VALU_MFMA_F32 x 2,048 flop) over the kernel's rocprof "
                    "duration, and SQ_VALU_MFMA_BUSY_CYCLES over the launch's SIMD-cycles, from the counter "
                    "runs of this build's sources (profiles/traffic.json); the kernel is bound by its culling"}
    # the same arm and cloud, 4,096 motions between consecutive configurations
    # (box validator 0.05 rad): the cloud variant of the edge kernel
    from plainmp_amd.ompl import BoxMotionValidator

    E = 4096
    S1, S2 = wl.Q[:E], wl.Q[E:2 * E]
    mv = BoxMotionValidator(cst, [0.05] * cst.q_dim())
    with torch.cuda.stream(stream):
        s1, s2 = _soa(torch, S1), _soa(torch, S2)
        ev = torch.empty(E, dtype=torch.uint8, device=dev)
        en = torch.empty(E, dtype=torch.int32, device=dev)
    stream.synchronize()
    _, tune_cloud = autotune.tune_edges(mv, cst, s1.data_ptr(), s2.data_ptr(), E, ev.data_ptr(), en.data_ptr(), sh,
                                        key="cloud_shelves_edges", repeats=3, cache=None)
    ms = _time_stream(torch, stream, lambda: mv._dev_check_motions(s1.data_ptr(), s2.data_ptr(), E, E, ev.data_ptr(),
                                                                   en.data_ptr(), sh), steps=10)
    checked = int(en.sum().item())
    res["cloud_shelves_edges"] = {"edges": E, "ms": ms, "edges_per_s": E / (ms * 1e-3),
                                  "waypoints_checked": checked, "configs_per_s": checked / (ms * 1e-3),
                                  "valid_frac": float(ev.float().mean().item()), "autotune": tune_cloud}
    # the cloud's spatial structure built on the device from device-resident
    # points (bounds, Morton sort, tiles, super-tiles): the shelf cloud and a
    # 1,048,576-point cloud
    from plainmp_amd._native import kdtree as kd

    builds = {}
    for name, pts in (("shelves_65536", wl.cloud), ("normal_1048576", np.random.default_rng(11).normal(
            size=(1 << 20, 3)))):
        n = len(pts)
        nt, ns, scr = kd.cloud_tile_sizes(n)
        with torch.cuda.stream(stream):
            dp = torch.tensor(np.ascontiguousarray(pts, dtype=np.float64), device=dev)
            bufs = [torch.empty(int(b), dtype=torch.uint8, device=dev)
                    for b in (48 * nt, 48 * nt, 1024 * nt, 2048 * nt, 32 * ns, scr)]
        args = [dp.data_ptr(), n] + [b.data_ptr() for b in bufs] + [sh]
        ms = _time_stream(torch, stream, lambda: kd._dev_cloud_build_tiles(*args), steps=10)
        builds[name] = {"points": n, "tiles": int(nt), "ms": ms, "points_per_s": n / (ms * 1e-3)}
    res["cloud_tile_build"] = builds
    # path utilities on the edges_mixed16 scene (host arrays in and out, so
    # these rates include the copies): every segment of a 4,096-state random
    # walk in one launch, and batched shortcutting of it (1,024 candidates a
    # round, 4 rounds)
    wl = synth.workload("edges_mixed16")
    cst = product_cst(wl)
    mv = BoxMotionValidator(cst, wl.box_width)
    rng = np.random.default_rng(11)
    P = wl.S1[0] + np.cumsum(rng.normal(scale=0.02, size=(4096, cst.q_dim())), axis=0)
    mv.check_path(P)
    t0 = time.perf_counter()
    for _ in range(5):
        seg = mv.check_path(P)
    t_path = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    S = mv.shortcut_path(P, 1024, 4, 0)
    t_short = time.perf_counter() - t0
    # batched multi-start position IK on the two-bar scene: 256 restarts x 60
    # iterations (link pose, link Jacobian and evaluate launches per iteration)
    from plainmp_amd.ik import solve_ik_batched

    wl_ik = synth.workload("plain_two_bars")
    cst_ik = product_cst(wl_ik)
    link = wl_ik.robot.sphere_specs[-1][0]
    target = [0.4, 0.2, 0.5]
    solve_ik_batched(cst_ik, link, target, wl_ik.robot.lower, wl_ik.robot.upper, n_restarts=256, seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ik = solve_ik_batched(cst_ik, link, target, wl_ik.robot.lower, wl_ik.robot.upper, n_restarts=256, seed=0)
    torch.cuda.synchronize()
    t_ik = time.perf_counter() - t0
    res["ik_batched_restarts"] = {"restarts": 256, "iterations": 60, "ms": t_ik * 1e3, "success": bool(ik.success),
                                  "first_success_restart": int(ik.n_trial),
                                  "restart_iterations_per_s": 256 * 60 / t_ik}
    res["_rooflines"] = rl
    res["path_check_and_shortcut"] = {"states": len(P), "check_path_ms": t_path * 1e3,
                                      "segments_per_s": (len(P) - 1) / t_path,
                                      "valid_segment_frac": float(seg.mean()), "shortcut_ms": t_short * 1e3,
                                      "shortcut_states_out": len(S), "candidates": 4 * 1024}
    return res


def _plain_per_call(wl, mv, n=200):
    """One edge per call: synchronous host->device copy, launch, device->host copy."""
    t0 = time.perf_counter()
    for e in range(n):
        mv.checkMotion(wl.S1[e], wl.S2[e])
    dt = time.perf_counter() - t0
    return {"edges": n, "s": dt, "edges_per_s": n / dt}


def _cpu_baseline(wl, budget_s=2.0, one_thread_s=1.0):
    """The FP64 oracle (a C++ restatement of the reference path) on host
    threads, and on one thread (the reference's own single-threaded call)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from support import oracle_cst

    oc = oracle_cst(wl)
    E = len(wl.S1)

    def sample(threads, budget, chunk):
        edges = checked = 0
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < budget:
            sl = slice((k * chunk) % E, (k * chunk) % E + chunk)
            ok, n = oc.check_motions(wl.S1[sl], wl.S2[sl], True, list(wl.box_width), 0.0, threads)
            edges += len(ok)
            checked += int(n.sum())
            k += 1
        return edges, checked, time.perf_counter() - t0

    threads = _host_threads()
    edges, checked, dt = sample(threads, budget_s, 1024)
    e1, c1, dt1 = sample(1, one_thread_s, 64)
    return {"value": checked / dt, "unit": "configs/s", "cores": threads, "kind": "port",
            "host": _host_info(), "edges_per_s": edges / dt,
            "sample": f"{edges} edges of edges_mixed16 ({checked} waypoints checked in reference order), "
                      f"{dt:.1f} s wall on {threads} threads",
            "one_thread": {"value": c1 / dt1, "unit": "configs/s", "edges_per_s": e1 / dt1,
                           "sample": f"{e1} edges ({c1} waypoints) in {dt1:.1f} s on 1 thread"}}


def _cpu_evaluate_baseline(budget_s=2.0):
    """Reference-semantics evaluate (value + Jacobian row) of grad_mixed32
    configurations by the FP64 oracle on host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from plainmp_amd import synth
    from support import oracle_cst

    threads = _host_threads()
    wl = synth.workload("grad_mixed32")
    oc = oracle_cst(wl)
    done = 0
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget_s:
        sl = slice((k * 2048) % len(wl.Q), (k * 2048) % len(wl.Q) + 2048)
        oc.evaluate_batch(wl.Q[sl], threads)
        done += len(wl.Q[sl])
        k += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "jacobians/s", "cores": threads, "kind": "port", "host": _host_info(),
            "sample": f"{done} grad_mixed32 configurations (evaluate: value + Jacobian row), {dt:.1f} s on {threads} threads"}


if __name__ == "__main__":
    main()
