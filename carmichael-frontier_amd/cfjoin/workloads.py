"""Benchmark and test workloads, built from the census rows in data/workloads.json.

  batch_instances("lambdaN", radii=(4, 5, 6, 7)) -- the batched-join workload: every ksmall
      instance of the 118 lambda-family moduli (census gpu/divisor_family_lambdaJ.json, bound
      U = N) with radius in `radii` and r <= rmax(D): 280 instances, 1.60e9 keys. The 111-modulus
      M0 family (U = N0) gives 268 instances, 2.15e9 keys.
  anchor("M0" | "lambdaN", r) -- the two flagship moduli (config-2 anchors).
Each instance carries the reference's logged (nD, nI, pairs) when a logged-counts table is given.
"""
from __future__ import annotations

import json
import os
from typing import Dict, List, Optional, Tuple

from . import instance as I


def logged_table(path: str) -> Dict[Tuple[str, int, int], dict]:
    with open(path) as f:
        rows = json.load(f)
    return {(r["family"], int(r["D"]), int(r["r"])): r for r in rows}


def batch_instances(family: str = "lambdaN", radii=(4, 5, 6, 7), logged: Optional[dict] = None,
                    limit: Optional[int] = None) -> List[I.Instance]:
    w = I.load_workloads()
    fam = w["families"][family]
    U = int(fam["U"])
    out = []
    for row in fam["rows"]:
        D, cap, rmax = int(row["D"]), int(row["cap"]), int(row["rmax"])
        rs = [r for r in radii if 1 <= r <= rmax]
        if not rs:
            continue
        P = I.pool(D, cap)
        for r in rs:
            inst = I.family_instance(D, cap, U, r, P=P)
            if inst is None:
                continue
            inst.tag = f"{family}:D{D}:r{r}"
            if logged is not None:
                inst.meta["logged"] = logged.get((family, D, r))
            out.append(inst)
            if limit and len(out) >= limit:
                return out
    return out


def anchor(family: str, r: int = 6, logged: Optional[dict] = None) -> I.Instance:
    inst = I.anchor_instance(family, r)
    if logged is not None:
        w = I.load_workloads()
        inst.meta["logged"] = logged.get((family, int(w["families"][family]["anchor"]), r))
    return inst


def serving_requests(n: int, seed: int = 0, logged: Optional[dict] = None, max_keys: float = 5e8,
                     radii=(4, 5, 6, 7, 8)) -> List[I.Instance]:
    """The serving workload: n modulus solves drawn uniformly (numpy default_rng(seed)) from the
    logged instances of both families with radius in `radii` and at most max_keys records (nD + nI),
    which mixes the batch class (divisor moduli, pools of 120-600 primes) with the larger anchors'
    class. Requires the logged-counts table (it sizes the draw)."""
    import numpy as np
    assert logged is not None, "serving_requests draws from the logged instances"
    w = I.load_workloads()
    cands = []
    for (fam, D, r), row in sorted(logged.items()):
        if r not in radii or row.get("path") not in ("dfilter", "host-exact") or row["nD"] + row["nI"] > max_keys:
            continue
        if row["nD"] + row["nI"] == 0:
            continue
        cands.append((fam, D, r))
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(cands), n)
    pools = {}
    out = []
    for k in picks:
        fam, D, r = cands[int(k)]
        f = w["families"][fam]
        row = next(x for x in f["rows"] if int(x["D"]) == D)
        if (fam, D) not in pools:
            pools[(fam, D)] = I.pool(D, int(row["cap"]))
        inst = I.family_instance(D, int(row["cap"]), int(f["U"]), r, P=pools[(fam, D)])
        inst.tag = f"{fam}:D{D}:r{r}"
        inst.meta["logged"] = logged.get((fam, D, r))
        out.append(inst)
    return out
