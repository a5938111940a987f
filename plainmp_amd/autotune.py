"""Autotuner of the launch schedule (pmp_schedule).

Two knobs decide how the persistent kernels map onto the MI355X:

* ``block`` / ``waves_per_cu`` of the edge kernel: workgroup size and how many
  persistent waves each CU holds (the grid never exceeds one wave per edge);
* ``eval_whole_chunks``: from how many 64-configuration chunks on an evaluate
  or trajectory batch walks each configuration as one unit instead of
  splitting it per sphere block and self pair;
* ``pair_shares``: into how many shares the pair-distance launch splits the
  primitives (grid z), each share walked by its own waves.

``tune_edges``, ``tune_evaluate`` and ``tune_pairs`` time every candidate on
the caller's own device buffers (HIP events around the edge kernel; stream
events around the evaluate and pair calls), keep the fastest, apply it to the constraint's context and
record it in a JSON cache keyed by device name and workload signature, so a
later run with the same workload applies the cached schedule without timing.
"""

from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

EDGE_BLOCKS = (64, 128, 256)  # the edge kernel is built for workgroups of at most 256
EDGE_WAVES_PER_CU = (4, 8, 12, 16, 20, 24, 32)
EVAL_WHOLE_CHUNKS = (16, 32, 64, 128, 256, 512, 1024, 4096)
PAIR_SHARES = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "plainmp_amd", "autotune.json")


def _device_name() -> str:
    import torch

    return torch.cuda.get_device_name(0)


def _load(path: str) -> Dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, table: Dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    except OSError:
        pass  # a read-only home: the schedule still applies to this process


class Schedule:
    """The context's current schedule; applies it with ``cst._dev_set_schedule``."""

    def __init__(self, block: int = 0, waves_per_cu: int = 0, eval_whole_chunks: int = 0, pair_shares: int = 0):
        self.block, self.waves_per_cu, self.eval_whole_chunks = block, waves_per_cu, eval_whole_chunks
        self.pair_shares = pair_shares

    def apply(self, cst) -> None:
        cst._dev_set_schedule(self.block, self.waves_per_cu, self.eval_whole_chunks, self.pair_shares)

    def replace(self, **kw) -> "Schedule":
        d = self.as_dict()
        d.update(kw)
        return Schedule(**d)

    def as_dict(self) -> Dict[str, int]:
        return {"block": self.block, "waves_per_cu": self.waves_per_cu, "eval_whole_chunks": self.eval_whole_chunks,
                "pair_shares": self.pair_shares}


def tune_edges(mv, cst, s1_ptr: int, s2_ptr: int, E: int, valid_ptr: int, nchk_ptr: int, stream: int,
               schedule: Optional[Schedule] = None, key: str = "", repeats: int = 5,
               blocks: Iterable[int] = EDGE_BLOCKS, waves: Iterable[int] = EDGE_WAVES_PER_CU,
               cache: Optional[str] = DEFAULT_CACHE) -> Tuple[Schedule, Dict]:
    """Pick block / waves_per_cu for the edge batch (device SoA s1, s2 of E
    edges); returns the applied schedule and a report."""
    schedule = schedule or Schedule()
    ckey = f"{_device_name()}|edges|{key}|E={E}"
    table = _load(cache) if cache else {}
    if ckey in table:
        schedule.block, schedule.waves_per_cu = table[ckey]["block"], table[ckey]["waves_per_cu"]
        schedule.apply(cst)
        return schedule, {"cached": True, **table[ckey]}
    results = []
    for b in blocks:
        for w in waves:
            schedule.replace(block=b, waves_per_cu=w).apply(cst)
            ms = []
            for _ in range(repeats + 1):
                mv._dev_check_motions(s1_ptr, s2_ptr, E, E, valid_ptr, nchk_ptr, stream)
                ms.append(mv._dev_edge_kernel_ms(stream))
            results.append((float(np.median(ms[1:])), b, w))
    results.sort()
    best_ms, schedule.block, schedule.waves_per_cu = results[0]
    schedule.apply(cst)
    rep = {"block": schedule.block, "waves_per_cu": schedule.waves_per_cu, "kernel_ms": best_ms,
           "candidates": len(results), "worst_ms": results[-1][0]}
    if cache:
        table[ckey] = rep
        _store(cache, table)
    return schedule, {"cached": False, **rep}


def tune_evaluate(cst, q_ptr: int, B: int, vals_ptr: int, jac_ptr: int, stream, schedule: Optional[Schedule] = None,
                  key: str = "", repeats: int = 5, choices: Iterable[int] = EVAL_WHOLE_CHUNKS,
                  cache: Optional[str] = DEFAULT_CACHE) -> Tuple[Schedule, Dict]:
    """Pick eval_whole_chunks for an evaluate batch of B configurations
    (``stream`` is a torch.cuda.Stream)."""
    import torch

    schedule = schedule or Schedule()
    ckey = f"{_device_name()}|evaluate|{key}|B={B}"
    table = _load(cache) if cache else {}
    if ckey in table:
        schedule.eval_whole_chunks = table[ckey]["eval_whole_chunks"]
        schedule.apply(cst)
        return schedule, {"cached": True, **table[ckey]}
    sh = stream.cuda_stream
    results = []
    for c in choices:
        schedule.replace(eval_whole_chunks=c).apply(cst)
        ms = []
        for _ in range(repeats + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            cst._dev_evaluate(q_ptr, B, B, vals_ptr, jac_ptr, sh)
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        results.append((float(np.median(ms[1:])), c))
    results.sort()
    best_ms, schedule.eval_whole_chunks = results[0]
    schedule.apply(cst)
    rep = {"eval_whole_chunks": schedule.eval_whole_chunks, "ms": best_ms, "candidates": len(results)}
    if cache:
        table[ckey] = rep
        _store(cache, table)
    return schedule, {"cached": False, **rep}


def _time_stream(stream, fn, repeats: int) -> float:
    """Median of `repeats` timed calls of fn on a torch.cuda.Stream (one
    untimed call first)."""
    import torch

    ms = []
    for _ in range(repeats + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    return float(np.median(ms[1:]))


def tune_pairs(cst, q_ptr: int, B: int, dist_ptr: int, grad_ptr: int, stream, schedule: Optional[Schedule] = None,
               key: str = "", repeats: int = 5, choices: Iterable[int] = PAIR_SHARES,
               cache: Optional[str] = DEFAULT_CACHE) -> Tuple[Schedule, Dict]:
    """Pick pair_shares for a pair-distance batch of B configurations
    (``grad_ptr`` 0: distances only; ``stream`` is a torch.cuda.Stream)."""
    schedule = schedule or Schedule()
    ckey = f"{_device_name()}|pairs|{key}|B={B}|grad={int(grad_ptr != 0)}"
    table = _load(cache) if cache else {}
    if ckey in table:
        schedule.pair_shares = table[ckey]["pair_shares"]
        schedule.apply(cst)
        return schedule, {"cached": True, **table[ckey]}
    sh = stream.cuda_stream
    results = []
    for c in choices:
        schedule.replace(pair_shares=c).apply(cst)
        results.append((_time_stream(stream, lambda: cst._dev_pair_clearance(q_ptr, B, B, dist_ptr, grad_ptr, sh),
                                     repeats), c))
    results.sort()
    best_ms, schedule.pair_shares = results[0]
    schedule.apply(cst)
    rep = {"pair_shares": schedule.pair_shares, "ms": best_ms, "candidates": len(results),
           "worst_ms": results[-1][0], "all": {str(c): round(t, 5) for t, c in sorted(results, key=lambda x: x[1])}}
    if cache:
        table[ckey] = rep
        _store(cache, table)
    return schedule, {"cached": False, **rep}


__all__ = ["Schedule", "tune_edges", "tune_evaluate", "tune_pairs", "EDGE_BLOCKS", "EDGE_WAVES_PER_CU",
           "EVAL_WHOLE_CHUNKS", "PAIR_SHARES"]
