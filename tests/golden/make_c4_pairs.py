#!/usr/bin/env python3
"""Writes tests/golden/c4_full_pairs.json: every pair of the large-radius instance -- lambda(N)
anchor, ksmall base, r = 8, 256-prime affordable pool, h <= 1, with the full D side (Dmin = 1, all
C(64,8) = 4 426 165 368 D subsets) -- found by the CPU oracle's swapped-role join (oracle.join_dsweep:
the I records hashed, the D subsets swept on 8 threads). Needs about 30 GB of host memory and a few
minutes. The GPU test compares the device solve's pair set with this file."""
import copy
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "carmichael-frontier_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from cfjoin import workloads as W  # noqa: E402


def main():
    logged = W.logged_table(os.path.join(HERE, "logged_instances.json"))
    inst = copy.deepcopy(W.anchor("lambdaN", 8, logged))
    inst.Dmin = 1
    t0 = time.time()
    res = O.join_dsweep(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                        inst.hmax, inst.rank_ids, threads=8)
    pairs = O.pair_set(res["pairs_d"], res["pairs_i"]).tolist()
    out = {"instance": "lambdaN anchor r=8, ksmall base, Dmin=1", "M": inst.M, "r": inst.r, "nD": res["nD"],
           "nI": res["nI"], "pairs": pairs, "oracle": "oracle.join_dsweep, threads=8",
           "seconds": {"i_hash": round(res["ti"], 1), "d_sweep": round(res["td"], 1)}}
    with open(os.path.join(HERE, "c4_full_pairs.json"), "w") as f:
        json.dump(out, f)
    print(f"nD={res['nD']} nI={res['nI']} pairs={len(pairs)} in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
