"""Exact join for instances whose insertion bound needs more than 128 bits.

The device trees carry each side's product as a 128-bit integer, so an insertion bound Icap above
2^128 cannot be expressed there. The reference routes such instances (and tiny ones) to a host join
in exact integers (gpu/mitm_gpu_mstar_driver.py:150-182, chosen at :207): D records from the ddump
program, the I side enumerated on the host with the product, cheapest-completion and stratum bounds,
joined by residue modulo M. Here the D records come from the device plan (cf_plan_create_d, the
ddump record set) and the I side is enumerated in Python integers with the dumper's bounds
(src/mitm64_idump_stream2.cpp:47-57: prod(I) < Icap, a first index whose cheapest completion
reaches Icap ends the loop, at most hmax indices at or above the split). The pair ids are the ones
the device path would give: D ids as the plan writes them, I ids as colex ranks (rank ids) or 8-bit
packed indices. Only instances beyond the device's 128-bit products belong here: their I sides are
small (the logged case, M0 at r = 14, has 23 records).
"""
from __future__ import annotations

from math import comb
from typing import Optional

import numpy as np

from . import lib as L
from . import pipeline


def solve(inst, ctx: Optional[L.Context] = None) -> pipeline.SolveResult:
    ctx = ctx or L.default_context()
    from . import xp
    r, M, INS = inst.r, inst.M, [int(q) for q in inst.INS]
    assert M < 1 << 64, "the host-exact join takes moduli below 2^64"
    i_mode, d_mode = pipeline.id_modes(inst)
    plan = L.Plan.deletion(inst.B0, r, M, inst.Dmin, d_mode, ctx=ctx)
    try:
        nD = plan.total()
        keys, ids = xp.empty(max(nD, 1)), xp.empty(max(nD, 1))
        if nD:
            plan.write(0, len(inst.B0), keys.ptr, ids.ptr)
    finally:
        plan.close()
    by_key = {}
    for k, v in zip(keys.get()[:nD].tolist(), ids.get()[:nD].tolist()):
        by_key.setdefault(k, []).append(v)
    top = inst.Icap - 1
    P0m = inst.P0 % M
    split, hmax = inst.split_idx, inst.hmax
    pd, pi = [], []
    nI = 0
    chosen = []

    def emit(prod):
        nonlocal nI
        nI += 1
        if i_mode == L.CF_ID_COLEX:
            iid = sum(comb(x, j + 1) for j, x in enumerate(chosen))
        else:
            iid = sum(x << (8 * j) for j, x in enumerate(chosen))
        for did in by_key.get(P0m * prod % M, ()):
            pd.append(did)
            pi.append(iid)

    def walk(lo, t, prod, large):
        if t == 0:
            emit(prod)
            return
        for i in range(lo, len(INS) - t + 1):
            cheapest = prod
            for q in INS[i:i + t]:
                cheapest *= q
            if cheapest > top:        # and every later first choice costs more
                break
            nl = large + (i >= split)
            if nl > hmax:
                continue
            chosen.append(i)
            walk(i + 1, t - 1, prod * INS[i], nl)
            chosen.pop()

    if inst.Icap > 1 and len(INS) >= r:
        walk(0, r, 1, 0)
    return pipeline.SolveResult(np.array(pd, np.uint64), np.array(pi, np.uint64), nD, nI,
                                {"path": "host-exact"})
