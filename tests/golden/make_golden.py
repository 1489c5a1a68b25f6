"""Writes the golden fixtures of tests/golden/ from the reference checkout (run once, where the
reference is present; the fixtures are committed, the reference is not needed afterwards).

  logged_instances.json  every instance line of the reference's run records
                         (gpu/records/<family>/D<modulus>.log): family, modulus, radius, stratum
                         bound, pool size, nD, nI, pairs, path (dfilter / host-exact / nothing).
  ref_streams.json       SHA-256 and counts of the record streams written by the compiled
                         reference dumpers (oracle/_ref, built by oracle/Makefile) on small
                         instances, and the MATCH sets of the compiled reference CPU join on the
                         two r = 6 anchors.
  small_records.npz      full record lists of a few tiny instances from the same tools.

usage: python tests/golden/make_golden.py [/path/to/reference] [--skip-anchors]
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "carmichael-frontier_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

LINE = re.compile(r"ksmall r=(\d+) h<=(\d+) \|INS\|=(\d+)(?:/(\d+))?( dfilter)?( rank)?( igpu)?(?: host-exact)?: "
                  r"D ([\d,]+)(?: [\d.]+s)? \| I ([\d,]+) rec [\d.]+s pairs=(\d+)")
NOTHING = re.compile(r"ksmall r=(\d+): affordable pool has (\d+) < r primes")


def parse_records(ref: str) -> list:
    out = []
    for fam_dir in sorted(os.listdir(os.path.join(ref, "gpu", "records"))):
        fam = fam_dir.split("_U")[0]
        d = os.path.join(ref, "gpu", "records", fam_dir)
        for fn in sorted(os.listdir(d)):
            D = int(fn[1:-4])
            for line in open(os.path.join(d, fn), errors="replace"):
                m = LINE.search(line)
                if m:
                    out.append(dict(family=fam, D=D, r=int(m.group(1)), hmax=int(m.group(2)), nINS=int(m.group(3)),
                                    path="host-exact" if "host-exact" in line else "dfilter",
                                    nD=int(m.group(8).replace(",", "")), nI=int(m.group(9).replace(",", "")),
                                    pairs=int(m.group(10))))
                    continue
                m = NOTHING.search(line)
                if m:
                    out.append(dict(family=fam, D=D, r=int(m.group(1)), path="nothing", nD=0, nI=0, pairs=0))
    return out


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def stream_fixtures(skip_anchors: bool) -> dict:
    import oracle as O
    from cfjoin import instance as I
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    w = I.load_workloads()
    cases = []
    fams = w["families"]

    def fam_inst(fam, D, r):
        row = next(x for x in fams[fam]["rows"] if int(x["D"]) == D)
        return I.family_instance(D, int(row["cap"]), int(fams[fam]["U"]), r)

    specs = [("M0", 1768248177696000, 3), ("M0", 1768248177696000, 4), ("M0", 1014485472000, 5),
             ("lambdaN", 10334428238534400, 6), ("lambdaN", 155016423578016000, 4)]
    for fam, D, r in specs:
        inst = fam_inst(fam, D, r)
        d = O.ref_ddump(inst.M, r, inst.Dmin, inst.B0)
        i = O.ref_idump(inst.M, r, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split, inst.hmax, inst.rank_ids, threads=1)
        cases.append(dict(kind="family", family=fam, D=D, r=r, nD=len(d), nI=len(i), d_sha=sha(d), i_sha=sha(i),
                          rank=inst.rank_ids))
    for seed, s in [(0, 10), (0, 100), (0, 10000), (1, 100), (2, 1000)]:
        inst = I.c1_instance(seed, s)
        d = O.ref_ddump(inst.M, 3, inst.Dmin, inst.B0) if len(inst.B0) == 64 else None
        i = O.ref_idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, None, None, False, threads=1)
        cases.append(dict(kind="c1", seed=seed, s=s, M=inst.M, nI=len(i), i_sha=sha(i)))
    anchors = []
    if not skip_anchors:
        for fam in ("M0", "lambdaN"):
            inst = I.anchor_instance(fam, 6)
            # Icap - 1: the reference CPU join admits prod(I) == Icap (src/mitm64.cpp:105)
            m, st = O.ref_mitm64(inst.M, 6, inst.P0 % inst.M, inst.Icap - 1, inst.Dmin, inst.B0, inst.INS, threads=8)
            anchors.append(dict(family=fam, r=6, matches=[[list(a), list(b)] for a, b in m], stats=st))
    return dict(cases=cases, anchors=anchors)


def small_records() -> dict:
    import oracle as O
    from cfjoin import instance as I
    out = {}
    inst = I.c1_instance(0, 10)
    i = O.ref_idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, None, None, False, threads=1)
    out["c1_seed0_s10_I"] = i
    w = I.load_workloads()
    row = next(x for x in w["families"]["M0"]["rows"] if int(x["D"]) == 1014485472000)
    inst = I.family_instance(1014485472000, int(row["cap"]), int(w["families"]["M0"]["U"]), 5)
    out["M0fam_D1014485472000_r5_D"] = O.ref_ddump(inst.M, 5, inst.Dmin, inst.B0)
    return out


def main():
    ref = next((a for a in sys.argv[1:] if not a.startswith("--")), "/root/reference")
    recs = parse_records(ref)
    with open(os.path.join(HERE, "logged_instances.json"), "w") as f:
        json.dump(recs, f, indent=0)
    print("logged instances:", len(recs))
    fx = stream_fixtures("--skip-anchors" in sys.argv)
    with open(os.path.join(HERE, "ref_streams.json"), "w") as f:
        json.dump(fx, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "small_records.npz"), **small_records())
    print("stream fixtures:", len(fx["cases"]), "anchors:", len(fx["anchors"]))


if __name__ == "__main__":
    main()
