"""Coverage audit: every logged instance of the two divisor families, solved on the device.

The reference certified its families by running every (census row, radius) instance and checking
the run logs for completeness (gpu/run_mstar_family.py, gpu/verify_family_coverage.py); each
instance's log line records its D and I record counts and its join matches
(gpu/mitm_gpu_mstar_driver.py:241,284: "D n | I n ... pairs=n"). This module rebuilds every logged
instance from the census (cfjoin/instance.py, the reference's recipe), solves them in batches
(cf_batch), and compares (nD, nI, pairs) with the log line by line; every pair is also re-checked
exactly on the host (residue identity and integrality of the completion, pipeline.recheck).

An instance whose insertion bound needs more than 128 bits is beyond the device's 128-bit products:
it is solved by the host-exact join (cfjoin/hostexact.py: D records from the device plan, I side in
Python integers), as the reference sends those to its host-exact join (gpu/mitm_gpu_mstar_driver.py:207),
and compared with its log line like the others.
"""
from __future__ import annotations

import json
import time
from typing import Dict, List, Optional

from . import hostexact
from . import instance as I
from . import lib as L
from . import pipeline

DEVICE_ICAP_BITS = 128


def logged_instances(path: str, families=("M0", "lambdaN")):
    """(log row, rebuilt instance or None) for every logged line of the given families."""
    with open(path) as f:
        rows = [r for r in json.load(f) if r["family"] in families]
    w = I.load_workloads()
    pools: Dict = {}
    out = []
    for r in rows:
        fam = w["families"][r["family"]]
        row = next(x for x in fam["rows"] if int(x["D"]) == int(r["D"]))
        key = (int(r["D"]), int(row["cap"]))
        if key not in pools:
            pools[key] = I.pool(*key)
        inst = I.family_instance(int(r["D"]), int(row["cap"]), int(fam["U"]), int(r["r"]), P=pools[key])
        out.append((r, inst))
    return out


def _batches(items, max_keys: float, max_insts: int):
    """Consecutive groups (in size order) of at most max_keys logged keys and max_insts instances,
    with the sum of their moduli below 2^64 - 1 (one batch is one key space)."""
    cur, keys, msum = [], 0, 0
    for it in sorted(items, key=lambda t: t[0]["nD"] + t[0]["nI"]):
        k = it[0]["nD"] + it[0]["nI"]
        m = it[1].M
        if cur and (keys + k > max_keys or len(cur) >= max_insts or msum + m >= (1 << 64) - 1):
            yield cur
            cur, keys, msum = [], 0, 0
        cur.append(it)
        keys += k
        msum += m
    if cur:
        yield cur


def audit(path: str, families=("M0", "lambdaN"), max_batch_keys: float = 2e9, max_batch_insts: int = 256,
          recheck: bool = True, ctx: Optional[L.Context] = None, log=None) -> dict:
    """Solves every logged instance; returns a report with per-instance mismatches (empty when the
    device reproduces the log)."""
    ctx = ctx or L.default_context()
    t0 = time.perf_counter()
    items = logged_instances(path, families)
    beyond = [(r, x) for r, x in items if x.Icap.bit_length() > DEVICE_ICAP_BITS]
    todo = [(r, x) for r, x in items if x.Icap.bit_length() <= DEVICE_ICAP_BITS]
    mismatches: List[dict] = []
    recheck_fail = 0
    pairs_total = keys_total = 0
    dev_ms = 0.0
    nbatch = 0
    for group in _batches(todo, max_batch_keys, max_batch_insts):
        insts = [x for _, x in group]
        b = pipeline.Batch(insts, ctx)
        try:
            res, st = b.solve()
        finally:
            b.close()
        nbatch += 1
        dev_ms += st["ms_total"]
        for (r, x), s in zip(group, res):
            keys_total += s.nD + s.nI
            pairs_total += len(s.pairs_d)
            got = (s.nD, s.nI, len(s.pairs_d))
            exp = (r["nD"], r["nI"], r["pairs"])
            if got != exp:
                mismatches.append(dict(family=r["family"], D=r["D"], r=r["r"], got=got, logged=exp))
            if recheck and len(s.pairs_d):
                recheck_fail += sum(not c["integral"] for c in pipeline.recheck(x, s))
        ctx.release_pool()
        if log:
            log(f"[audit] batch {nbatch}: {len(group)} instances, {sum(s.nD + s.nI for s in res):,} keys, "
                f"{st['ms_total']:.1f} ms device")
    for r, x in beyond:
        s = hostexact.solve(x, ctx)
        keys_total += s.nD + s.nI
        pairs_total += len(s.pairs_d)
        got = (s.nD, s.nI, len(s.pairs_d))
        if got != (r["nD"], r["nI"], r["pairs"]):
            mismatches.append(dict(family=r["family"], D=r["D"], r=r["r"], got=got,
                                   logged=(r["nD"], r["nI"], r["pairs"]), path="host-exact"))
        if recheck and len(s.pairs_d):
            recheck_fail += sum(not c["integral"] for c in pipeline.recheck(x, s))
        if log:
            log(f"[audit] host-exact: {r['family']} D={r['D']} r={r['r']} (Icap {x.Icap.bit_length()} bits): "
                f"D {s.nD:,} | I {s.nI:,} pairs={len(s.pairs_d)}")
    return dict(instances=len(items), solved=len(todo), solved_host_exact=len(beyond), batches=nbatch,
                keys=keys_total, pairs=pairs_total,
                device_ms=dev_ms, wall_s=time.perf_counter() - t0, mismatches=mismatches,
                recheck_failures=recheck_fail,
                beyond_device=[dict(family=r["family"], D=r["D"], r=r["r"], icap_bits=x.Icap.bit_length(),
                                    logged=(r["nD"], r["nI"], r["pairs"])) for r, x in beyond])
