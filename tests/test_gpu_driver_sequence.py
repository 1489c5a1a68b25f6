"""The mstar driver's GPU sequence, restated through the drop-in modules, against logged lines.

gpu/mitm_gpu_mstar_driver.py runs one instance with DFILTER=1 and IGEN=gpu as follows:
  * D side (build_D, :101-115): the ddump program's records read from its stdout, split into
    chunks by numpy, each chunk moved to the device (cp.asarray) and sorted (sort_side), host
    arrays dropped and the pool released (cp.get_default_memory_pool().free_all_blocks());
  * I side (:213-238): igen_gpu.generate_chunks, the total read from StopIteration.value; each chunk
    sorted and joined against every D chunk (join_sorted); hits copied back (cp.asnumpy) and decoded
    (unrank_colex for the colex I ids, decode_ids for the 6-bit D ids); the pool released after
    every chunk; cp.cuda.Device().mem_info read for the log line.
The test below performs that sequence with `cfjoin.xp` imported as `cp` and the ddump stand-in
(bin/mitm64_ddump) run as a child process over the same pipe protocol, and compares the record
counts and pair count with the reference's log line and the pair set with the fused solve.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from cfjoin import instance as I
from cfjoin import pipeline
from cfjoin import verify as V

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DDUMP = os.path.join(ROOT, "carmichael-frontier_amd", "bin", "mitm64_ddump")
LOGGED = os.path.join(ROOT, "tests", "golden", "logged_instances.json")


def say(*a):
    print("CASE", *a, flush=True)


def logged_line(family, D, r):
    with open(LOGGED) as f:
        return next(x for x in json.load(f) if x["family"] == family and x["D"] == D and x["r"] == r)


def d_side_through_pipe(cp, G, B0, M, r, Dmin, per_chunk):
    """The ddump records over the pipe, chunked, on the device and sorted."""
    text = f"{M} {r} {Dmin}\n64 " + " ".join(str(p) for p in sorted(B0)) + "\n"
    pr = subprocess.run([sys.executable, DDUMP], input=text.encode(), capture_output=True, check=True)
    recs = np.frombuffer(pr.stdout, dtype=np.uint64).reshape(-1, 2)
    total = [int(l.split()[1]) for l in pr.stderr.decode().splitlines() if l.startswith("TOTAL")]
    assert total == [recs.shape[0]]
    chunks = []
    for rows in np.array_split(np.arange(recs.shape[0]), max(1, min(16, recs.shape[0] // per_chunk + 1))):
        if rows.size:
            k = cp.asarray(np.ascontiguousarray(recs[rows, 0]))
            i = cp.asarray(np.ascontiguousarray(recs[rows, 1]))
            chunks.append(G.sort_side(k, i))
            del k, i
    del recs
    cp.get_default_memory_pool().free_all_blocks()
    return chunks, total[0]


@pytest.mark.parametrize("family,D,r,per_chunk,i_chunk", [
    ("M0", 1768248177696000, 5, 1_000_000, 20_000_000),      # 3 D chunks, 3 I chunks
    ("M0", 1768248177696000, 6, 50_000_000, 200_000_000),    # the driver's own chunk sizes: 1 D, 2 I chunks
    ("lambdaN", 155016423578016000, 6, 50_000_000, 200_000_000),
])
def test_driver_dfilter_igen_sequence(workloads, family, D, r, per_chunk, i_chunk):
    import igen_gpu as IG
    import mitm_gpu2 as G
    from cfjoin import xp as cp

    line = logged_line(family, D, r)
    fam = workloads["families"][family]
    row = next(x for x in fam["rows"] if int(x["D"]) == D)
    inst = I.family_instance(D, int(row["cap"]), int(fam["U"]), r)
    assert len(inst.INS) == line["nINS"] and inst.hmax == line["hmax"]
    say(f"driver sequence {family} D={D} r={r}: logged nD={line['nD']} nI={line['nI']} pairs={line['pairs']}")
    Dch, nD = d_side_through_pipe(cp, G, inst.B0, inst.M, r, inst.Dmin, per_chunk)
    gen = IG.generate_chunks(inst.INS, r, inst.M, inst.P0, inst.Icap, I.SPLIT_DEFAULT, inst.hmax, max_records=i_chunk)
    nI_seen, n_chunks, hits = 0, 0, []
    while True:
        try:
            Ik, Iid = next(gen)
        except StopIteration as stop:
            nI = stop.value
            break
        nI_seen += int(Ik.size)
        n_chunks += 1
        Ik_s, Iid_s = G.sort_side(Ik, Iid)
        del Ik, Iid
        for Dk_s, Did_s in Dch:
            pd, pi = G.join_sorted(Dk_s, Did_s, Ik_s, Iid_s)
            if pd.size == 0:
                continue
            i_idx = [V.unrank_colex(v, r, len(inst.INS)) for v in cp.asnumpy(pi)]
            d_idx = G.decode_ids(pd, r, 6)
            for d, i in zip(d_idx, i_idx):
                V.completion(inst.B0, inst.INS, d, i, inst.P0, inst.M)   # raises on unequal residues
                hits.append((tuple(sorted(inst.B0[x] for x in d)), tuple(inst.INS[x] for x in i)))
        del Ik_s, Iid_s
        cp.get_default_memory_pool().free_all_blocks()
    free_b, total_b = cp.cuda.Device().mem_info
    assert 0 < free_b <= total_b
    assert (nD, nI, nI_seen, len(hits)) == (line["nD"], line["nI"], line["nI"], line["pairs"])
    assert n_chunks >= (2 if nI > i_chunk else 1)     # the multi-chunk path is exercised where it exists
    # the same pair set as the fused solve of the instance
    res = pipeline.solve(inst)
    d_dec, i_dec = pipeline.decoders(inst)
    fused = sorted((tuple(sorted(inst.B0[x] for x in d_dec(int(a)))), tuple(inst.INS[x] for x in i_dec(int(b))))
                   for a, b in zip(res.pairs_d, res.pairs_i))
    assert sorted(hits) == fused
