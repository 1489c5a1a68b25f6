#!/usr/bin/env python3
"""Calibrates the CPU baseline: the oracle's join (oracle.c or_cpu_join) against the reference's
own CPU join (src/mitm64.cpp, compiled into oracle/_ref/mitm64 by `make -C oracle ref`) on
identical inputs.

TEST INFRASTRUCTURE ONLY (it runs the checker and the compiled reference, never the product).

For the two r = 6 anchors (M0 = 1768248177696000 with U = N0, lambda(N) = 155016423578016000
with U = N; ksmall base, full capped pool) both programs get the same modulus, base, pool, thread
count, inclusive bound Icap - 1 and Dmin (src/mitm64.cpp:126-134); the product bound already
implies the anchors' h-stratum, so the reference's I leaves equal the logged record count. Each
runs `runs` times; the time is td + ti from each program's own clock (src/mitm64.cpp:156-159), and
keys/s = (D leaves + I leaves) / (td + ti), median over the runs. The reference keeps one D entry
per residue (the largest product), so its MATCH set must be contained in the oracle's pair set.

    python oracle/calibrate_cpu.py [--threads T] [--runs 3] [--out oracle/calibration.json]
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "carmichael-frontier_amd")]

import oracle as O  # noqa: E402
from cfjoin import instance as I  # noqa: E402
from cfjoin import pipeline  # noqa: E402


def host_cpu() -> dict:
    """Model, logical CPUs, physical cores and SMT state of this host (from /proc and /sys)."""
    info = {"model": platform.processor() or "unknown", "logical_cpus": os.cpu_count()}
    try:
        cores, model = set(), None
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
        info["model"] = model or info["model"]
        info["physical_cores_visible"] = len(cores) or None
    except OSError:
        pass
    try:
        info["smt"] = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        info["smt"] = None
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def calibrate(which: str, threads: int, runs: int) -> dict:
    inst = I.anchor_instance(which, 6)
    args = (inst.M, inst.r, inst.P0 % inst.M)
    ref, port = [], []
    ref_matches = None
    port_pairs = None
    for _ in range(runs):
        m, st = O.ref_mitm64(*args, inst.Icap - 1, inst.Dmin, inst.B0, inst.INS, threads)
        ref.append(st)
        ref_matches = set(m)
        res = O.cpu_join(*args, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx, inst.hmax, inst.rank_ids,
                         threads=threads)
        port.append(res)
        d_dec, i_dec = pipeline.decoders(inst)
        port_pairs = {(tuple(sorted(inst.B0[x] for x in d_dec(int(a)))), tuple(sorted(inst.INS[x] for x in i_dec(int(b)))))
                      for a, b in zip(res["pairs_d"], res["pairs_i"])}
    keys_ref = ref[0]["D_leaves"] + ref[0]["I_leaves"]
    keys_port = port[0]["nD"] + port[0]["nI"]
    t_ref = statistics.median(s["td"] + s["ti"] for s in ref)
    t_port = statistics.median(p["td"] + p["ti"] for p in port)
    return {
        "instance": f"{which} r=6 (M={inst.M}, |INS|={len(inst.INS)})",
        "reference_mitm64": {"D_leaves": ref[0]["D_leaves"], "I_leaves": ref[0]["I_leaves"],
                             "matches": ref[0]["matches"], "td_ti_s": [round(s["td"] + s["ti"], 3) for s in ref],
                             "keys_per_s": keys_ref / t_ref},
        "port_or_cpu_join": {"nD": port[0]["nD"], "nI": port[0]["nI"], "pairs": len(port[0]["pairs_d"]),
                             "td_ti_s": [round(p["td"] + p["ti"], 3) for p in port], "keys_per_s": keys_port / t_port},
        "same_records": keys_ref == keys_port,
        "reference_matches_within_port_pairs": bool(ref_matches <= port_pairs),
        "port_over_reference_speed": (keys_port / t_port) / (keys_ref / t_ref),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "calibration.json"))
    a = ap.parse_args()
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref, where the reference sources are present)")
    t0 = time.time()
    out = {"host_cpu": host_cpu(), "threads": a.threads, "runs": a.runs,
           "anchors": [calibrate(w, a.threads, a.runs) for w in ("M0", "lambdaN")]}
    out["port_over_reference_speed"] = statistics.median(x["port_over_reference_speed"] for x in out["anchors"])
    out["wall_s"] = round(time.time() - t0, 1)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
