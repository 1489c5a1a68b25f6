#!/usr/bin/env python3
"""Writes tests/golden/c2_pairs.json: for the seeded k = 64, r = 6 moduli (cfjoin.instance.c2_instance,
seeds 0-2, bound s = 1000, 6e8-1.7e9 records each) the record counts and every pair, found by the CPU
oracle's join (oracle.cpu_join, the reference CPU join's algorithm, 8 threads). The GPU test solves the
same instances and compares. Takes a few minutes."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "carmichael-frontier_amd"), os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from cfjoin import instance as I  # noqa: E402


def main():
    out = []
    for seed in (0, 1, 2):
        inst = I.c2_instance(seed, 1000)
        t0 = time.time()
        res = O.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                         inst.hmax, inst.rank_ids, threads=8)
        out.append({"seed": seed, "s": 1000, "M": inst.M, "nINS": len(inst.INS), "nD": res["nD"], "nI": res["nI"],
                    "pairs": O.pair_set(res["pairs_d"], res["pairs_i"]).tolist()})
        print(f"seed {seed}: M={inst.M} nD={res['nD']} nI={res['nI']} pairs={len(res['pairs_d'])} "
              f"in {time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "c2_pairs.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
