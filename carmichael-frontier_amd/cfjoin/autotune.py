"""Autotuner for batched solves: picks the probe-table slice per partition, the D records per
bucket, the fill of the I capacity segments and the probe grid (cf_tuning) by timing candidate
settings on the prepared batch itself.

Search: coordinate descent from the defaults -- one knob at a time over its candidate values, the
others held at their best so far, each candidate timed as the best of `reps` solves. Every candidate's
result is compared with the default setting's (per-instance record counts and pair sets): a knob
may only change speed, so a difference is an error, not a choice.

With graph=True each candidate is timed as the captured solve the batch will run (cf_batch_capture
on a copy of the batch, then replays): in a replay the radius groups run side by side and share the
L2 slices, so the best eager setting is not always the best replayed one.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import lib as L

DEFAULTS = dict(slice_kib=2048, d_per_bucket=2, seg_fill_pct=70, probe_blocks=256)
CANDIDATES = {
    "slice_kib": [1024, 2048, 4096],
    "d_per_bucket": [1, 2, 3],
    "seg_fill_pct": [60, 70, 80],
    "probe_blocks": [],    # one 1024-thread probe block per CU: its filter slice fills the LDS
}


def _signature(results) -> List[Tuple[int, int, bytes]]:
    return [(r.nD, r.nI, r.pair_array().tobytes()) for r in results]


def _time(batch, setting: Dict[str, int], reps: int, ctx: L.Context):
    batch.set_tuning(**setting)
    res, _ = batch.solve()                 # warm: allocations settle in the pool
    best = float("inf")
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        batch.solve()
        ctx.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best * 1e3, res


def _time_graph(batch, setting: Dict[str, int], reps: int, ctx: L.Context):
    from . import pipeline
    b = pipeline.Batch(batch.insts, ctx)
    try:
        b.set_tuning(**setting)
        b.capture()
        res, _ = b.solve()
        best = float("inf")
        for _ in range(reps):
            ctx.synchronize()
            t0 = time.perf_counter()
            b.solve()
            ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
    finally:
        b.close()
    return best * 1e3, res


def tune(batch, reps: int = 2, candidates: Optional[Dict[str, List[int]]] = None, log=None,
         graph: bool = False) -> Dict:
    """Tune a prepared, not yet captured pipeline.Batch in place; returns
    {"best": setting, "ms": best time, "default_ms": ..., "trials": [(setting, ms), ...]}.
    graph: time each candidate as a captured solve (replays) rather than as an eager one."""
    cands = candidates or CANDIDATES
    ctx = batch.ctx
    timer = _time_graph if graph else _time
    best = dict(DEFAULTS)
    best_ms, ref = timer(batch, best, reps, ctx)
    ref_sig = _signature(ref)
    default_ms = best_ms
    trials = [(dict(best), best_ms)]
    for knob, values in cands.items():
        for v in values:
            if v == best[knob]:
                continue
            setting = dict(best, **{knob: v})
            ms, res = timer(batch, setting, reps, ctx)
            if _signature(res) != ref_sig:
                raise AssertionError(f"tuning {setting} changed the result of the solve")
            trials.append((setting, ms))
            if log:
                log(f"[tune] {setting} {ms:.2f} ms")
            if ms < best_ms:
                best, best_ms = setting, ms
    batch.set_tuning(**best)
    return {"best": best, "ms": best_ms, "default_ms": default_ms, "trials": trials}
