#!/usr/bin/env python3
"""Writes tests/golden/posctl_pairs.json: the positive controls' pair sets from the CPU oracle.

For each control (cfjoin.instance.control_instance: an incumbent's own modulus, the ksmall base and
every pool prime below 4000 outside it, bound U = n + 1) the record counts and every pair of the
oracle's join (oracle.cpu_join, the reference CPU join's algorithm, 8 threads), with the smallest
integral completion among them. For the validator's drop-in composition (nins = 30, full D side,
unpruned I side; gpu/validate_mitm_gpu2_pos.py:50-58) every equal-residue pair as sorted prime
tuples, from the oracle's join with the roles swapped (oracle.join_dsweep, Dmin = 1, Icap = the
largest product + 1). Takes a few minutes."""
import json
import os
import sys
import time
from math import prod

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "carmichael-frontier_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from cfjoin import instance as I  # noqa: E402
from cfjoin import pipeline  # noqa: E402
from cfjoin import verify as V  # noqa: E402


def min_completion(inst, pd, pi):
    d_dec, i_dec = pipeline.decoders(inst)
    best = None
    for dv, iv in zip(pd, pi):
        c = V.completion(inst.B0, inst.INS, d_dec(int(dv)), i_dec(int(iv)), inst.P0, inst.M)
        if c is not None and (best is None or c[1] < best):
            best = c[1]
    return best


def main():
    out = {"solve": {}, "compose": {}}
    for name in I.CONTROLS:
        c = I.control_instance(name)
        inst = c.inst
        t0 = time.time()
        res = O.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                         inst.hmax, inst.rank_ids, threads=8)
        best = min_completion(inst, res["pairs_d"], res["pairs_i"])
        out["solve"][name] = {"M": inst.M, "r": inst.r, "nINS": len(inst.INS), "nD": res["nD"], "nI": res["nI"],
                              "min_completion": str(best),
                              "pairs": O.pair_set(res["pairs_d"], res["pairs_i"]).tolist()}
        print(f"{name}: r={inst.r} |INS|={len(inst.INS)} nD={res['nD']} nI={res['nI']} pairs={len(res['pairs_d'])} "
              f"min completion {'= n' if best == c.N else best} in {time.time() - t0:.1f} s", flush=True)
    for name in ("N", "N_prev"):
        c = I.control_instance(name, nins=30)
        inst = c.inst
        t0 = time.time()
        icap_all = prod(inst.INS[-inst.r:]) + 1
        res = O.join_dsweep(inst.M, inst.r, inst.P0 % inst.M, icap_all, 1, inst.B0, inst.INS, None, None, False,
                            threads=8)
        d_dec, _ = pipeline.decoders(inst)
        pairs = sorted([sorted(inst.B0[x] for x in d_dec(int(dv))), [inst.INS[x] for x in V.decode_packed(int(iv), inst.r, 8)]]
                       for dv, iv in zip(res["pairs_d"], res["pairs_i"]))
        out["compose"][name] = {"M": inst.M, "r": inst.r, "INS": inst.INS, "nD": res["nD"], "nI": res["nI"],
                                "pairs": pairs}
        print(f"{name} composition: nD={res['nD']} nI={res['nI']} pairs={len(pairs)} in {time.time() - t0:.1f} s",
              flush=True)
    with open(os.path.join(HERE, "posctl_pairs.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
