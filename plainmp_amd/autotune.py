"""Autotuner of the launch schedule (pmp_schedule).

Two knobs decide how the persistent kernels map onto the MI355X:

* ``block`` / ``waves_per_cu`` of the edge kernel: workgroup size and how many
  persistent waves each CU holds (the grid never exceeds one wave per edge);
* ``eval_whole_chunks``: from how many 64-configuration chunks on an evaluate
  or trajectory batch walks each configuration as one unit instead of
  splitting it per sphere block and self pair.

``tune_edges`` and ``tune_evaluate`` time every candidate on the caller's own
device buffers (HIP events around the edge kernel; stream events around the
evaluate call), keep the fastest, apply it to the constraint's context and
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

    def __init__(self, block: int = 0, waves_per_cu: int = 0, eval_whole_chunks: int = 0):
        self.block, self.waves_per_cu, self.eval_whole_chunks = block, waves_per_cu, eval_whole_chunks

    def apply(self, cst) -> None:
        cst._dev_set_schedule(self.block, self.waves_per_cu, self.eval_whole_chunks)

    def as_dict(self) -> Dict[str, int]:
        return {"block": self.block, "waves_per_cu": self.waves_per_cu, "eval_whole_chunks": self.eval_whole_chunks}


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
            Schedule(b, w, schedule.eval_whole_chunks).apply(cst)
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
        Schedule(schedule.block, schedule.waves_per_cu, c).apply(cst)
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


__all__ = ["Schedule", "tune_edges", "tune_evaluate", "EDGE_BLOCKS", "EDGE_WAVES_PER_CU", "EVAL_WHOLE_CHUNKS"]
