"""Serving loop over the C ABI's continuous-batching server (cf_server_*): requests are submitted
as they arrive, a free worker solves everything that has arrived as one batch (admitted by its
priced memory footprint), and the workers' streams overlap on the GPU.

  Server(workers, max_batch, mem_budget)  -- submit(inst, tag), next(timeout_ms), close()
  poisson_arrivals(n, rate, seed)         -- arrival times of a Poisson stream
  run_stream(server, insts, arrivals)     -- submit on schedule, collect every result
  run_one_at_a_time(insts)                -- each request solved alone, the next submitted only
                                             when the previous one finished (the comparison run)
  summary(served, wall_s)                 -- p50 / p99 latency and keys/s
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import lib as L
from . import pipeline


@dataclass
class Served:
    tag: int
    status: int
    error: str
    batch_size: int
    nD: int
    nI: int
    pairs_d: np.ndarray
    pairs_i: np.ndarray
    t_submit_us: float
    t_start_us: float
    t_done_us: float

    @property
    def latency_ms(self) -> float:
        return (self.t_done_us - self.t_submit_us) / 1e3

    def pair_array(self) -> np.ndarray:
        return pipeline.SolveResult(self.pairs_d, self.pairs_i, self.nD, self.nI).pair_array()


class Server:
    def __init__(self, workers: int = 4, max_batch: int = 64, mem_budget: int = 0, device: int = 0):
        h = ctypes.c_void_p()
        L.check(L.load().cf_server_create(device, workers, max_batch, mem_budget, ctypes.byref(h)), "cf_server_create")
        self.handle = h
        self._keep = {}

    def submit(self, inst, tag: int) -> None:
        keep: list = []
        ci = pipeline.make_cinstance(inst, keep)
        L.check(L.load().cf_server_submit(self.handle, ctypes.byref(ci), tag), "cf_server_submit")

    def next(self, timeout_ms: int = 1000) -> Optional[Served]:
        out = L.cf_served()
        rc = L.load().cf_server_next(self.handle, timeout_ms, ctypes.byref(out))
        L.check(rc if rc < 0 else 0, "cf_server_next")
        if rc == 0:
            return None
        n = int(out.npairs)
        pd = np.ctypeslib.as_array(out.pairs_d, shape=(n,)).copy() if n else np.zeros(0, np.uint64)
        pi = np.ctypeslib.as_array(out.pairs_i, shape=(n,)).copy() if n else np.zeros(0, np.uint64)
        s = Served(int(out.tag), int(out.status), out.error.decode(errors="replace"), int(out.batch_size), int(out.nD),
                   int(out.nI), pd, pi, out.t_submit_us, out.t_start_us, out.t_done_us)
        L.check(L.load().cf_server_release(self.handle, out.tag), "cf_server_release")
        return s

    def close(self) -> None:
        if getattr(self, "handle", None):
            L.check(L.load().cf_server_destroy(self.handle), "cf_server_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def run_stream(server: Server, insts, arrivals) -> List[Served]:
    """Submit insts[k] at time arrivals[k] (seconds from now), collect all results (in finish order)."""
    t0 = time.perf_counter()
    out: List[Served] = []
    k = 0
    n = len(insts)
    while len(out) < n:
        now = time.perf_counter() - t0
        while k < n and arrivals[k] <= now:
            server.submit(insts[k], k)
            k += 1
        wait_ms = int(max(0.0, (arrivals[k] - (time.perf_counter() - t0)) * 1e3)) if k < n else 200
        r = server.next(min(wait_ms, 200))
        if r is not None:
            if r.status:
                raise RuntimeError(f"request {r.tag} failed: {r.error}")
            out.append(r)
    return out


def run_one_at_a_time(insts, workers: int = 1) -> List[Served]:
    """Each request alone: submitted only after the previous one finished (no batching, no overlap)."""
    server = Server(workers=workers, max_batch=1)
    out = []
    try:
        for k, inst in enumerate(insts):
            server.submit(inst, k)
            r = None
            while r is None:
                r = server.next(1000)
            if r.status:
                raise RuntimeError(f"request {r.tag} failed: {r.error}")
            out.append(r)
    finally:
        server.close()
    return out


def summary(served: List[Served], wall_s: float) -> dict:
    lat = np.array([s.latency_ms for s in served])
    keys = sum(s.nD + s.nI for s in served)
    return {"requests": len(served), "keys": int(keys), "wall_s": round(wall_s, 4),
            "keys_per_s": keys / wall_s if wall_s > 0 else None,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "mean_batch": float(np.mean([s.batch_size for s in served]))}
