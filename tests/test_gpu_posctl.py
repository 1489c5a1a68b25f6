"""Positive controls on the GPU: instances that must contain a known k = 64 Carmichael number.

Each control is an incumbent's own modulus with the ksmall base and an insertion pool holding the
incumbent's inserted primes (cfjoin.instance.control_instance; gpu/validate_mitm_gpu2_pos.py:20-36).
The join must emit the pair (D_need, I_need) whose completion is the incumbent, the exact re-check
must accept it as Carmichael with 64 prime factors, and the pair set must equal the CPU oracle's
(tests/golden/posctl_pairs.json, written by tests/golden/make_posctl.py).

Two routes:
  * the fused solve (cf_solve) over every pool prime below 4000 outside the base, radii 6 to 9;
  * the drop-in composition the reference's validator runs (gpu/validate_mitm_gpu2_pos.py:50-73):
    the D side in 4 rank chunks joined by xp.concatenate, checked against one shot with xp.sort
    and xp.all, the unpruned I side, sort_side, join_sorted and decode_ids.
"""
import json
import os
from math import comb, prod

import numpy as np
import pytest

from cfjoin import instance as I
from cfjoin import pipeline
from cfjoin import verify as V

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "posctl_pairs.json")


def say(*a):
    print("CASE", *a, flush=True)


@pytest.fixture(scope="module")
def posctl():
    with open(GOLDEN) as f:
        return json.load(f)


def completions(inst, d_tuples, i_tuples):
    """Exact completion n of every integral pair (V.completion raises on a residue mismatch)."""
    out = []
    for d, i in zip(d_tuples, i_tuples):
        c = V.completion(inst.B0, inst.INS, d, i, inst.P0, inst.M)
        if c is not None:
            out.append((c[1], c[0], [inst.B0[x] for x in d], [inst.INS[x] for x in i]))
    return out


@pytest.mark.parametrize("name", list(I.CONTROLS))
def test_control_solve_recovers_incumbent(name, posctl):
    c = I.control_instance(name)
    inst = c.inst
    fx = posctl["solve"][name]
    say(f"control {name}: M={inst.M} r={inst.r} |INS|={len(inst.INS)} expecting nD={fx['nD']} nI={fx['nI']}")
    res = pipeline.solve(inst)
    assert (res.nD, res.nI) == (fx["nD"], fx["nI"])
    assert np.array_equal(res.pair_array(), np.array(fx["pairs"], np.uint64).reshape(-1, 2))
    d_dec, i_dec = pipeline.decoders(inst)
    found = completions(inst, [d_dec(int(v)) for v in res.pairs_d], [i_dec(int(v)) for v in res.pairs_i])
    assert any(sorted(D) == c.D_need and sorted(Iq) == c.I_need for _, _, D, Iq in found), "(D_need, I_need) missing"
    assert c.N in [x[0] for x in found]
    # the smallest completion is the oracle's; for the three 148-digit incumbents it is the incumbent
    # itself, the 149-digit one has a smaller completion at its own radius
    best_n, best_fac, _, _ = min(found)
    assert best_n == int(fx["min_completion"]) <= c.N
    assert best_n == c.N or name == "N149"
    ok, n, _ = V.verify_certificate(best_fac)
    assert ok and n == best_n and len(best_fac) == 64
    # every pair the re-check accepts is a genuine Carmichael number of the reference's checker's kind
    checks = pipeline.recheck(inst, res)
    assert any(x.get("carmichael") and x["n_digits"] == len(str(c.N)) for x in checks)
    # every integral completion is Carmichael: n = 1 (mod M) and p - 1 divides M for every p
    assert all(x["carmichael"] for x in checks if x["integral"])


@pytest.mark.parametrize("name", ["N_prev", "N"])
def test_control_dropin_composition(name, posctl):
    import mitm_gpu2 as G
    from cfjoin import xp as cp

    c = I.control_instance(name, nins=30)
    inst = c.inst
    R, M, B0, INS = inst.r, inst.M, inst.B0, inst.INS
    fx = posctl["compose"][name]
    assert INS == fx["INS"]
    totD = comb(len(B0), R)
    say(f"composition {name}: M={M} r={R} D side C(64,{R})={totD} in 4 rank chunks, I side C(30,{R})")
    cuts = [totD * q // 4 for q in range(5)]
    parts = [G.generate(B0, R, M, 1, bits=6, lo=a, hi=b) for a, b in zip(cuts, cuts[1:])]
    Dk = cp.concatenate([p[0] for p in parts])
    Did = cp.concatenate([p[1] for p in parts])
    del parts
    k1, _ = G.generate(B0, R, M, 1, bits=6)
    assert Dk.size == totD and cp.all(cp.sort(Dk) == cp.sort(k1)), "chunk union != single shot"
    del k1
    Ik, Iid = G.generate(INS, R, M, inst.P0 % M, bits=8)
    assert Ik.size == comb(len(INS), R)
    Dk_s, Did_s = G.sort_side(Dk, Did)
    Ik_s, Iid_s = G.sort_side(Ik, Iid)
    del Dk, Did, Ik, Iid
    pd, pi = G.join_sorted(Dk_s, Did_s, Ik_s, Iid_s)
    d_t, i_t = G.decode_ids(pd, R, 6), G.decode_ids(pi, R, 8)
    assert len(d_t) == pd.size == pi.size
    pairs = sorted([sorted(B0[x] for x in d), sorted(INS[x] for x in i)] for d, i in zip(d_t, i_t))
    assert pairs == fx["pairs"]
    # every pair has exactly equal residues; the validator keeps those under its own product bound
    found = completions(inst, d_t, i_t)
    ptb = prod(B0[-R:])
    icap_v = max(1, min(prod(INS[-R:]), (c.N * ptb) // inst.P0 + 1))
    kept = [x for x in found if prod(x[3]) < icap_v]
    assert c.N in [x[0] for x in kept], "the incumbent is not recovered"
    best_n, best_fac, _, _ = min(kept)
    assert best_n == c.N and V.verify_certificate(best_fac)[0]
    cp.get_default_memory_pool().free_all_blocks()


@pytest.mark.parametrize("mode", ["partition", "partition_dglobal"])
def test_control_full_d_side_in_range_tables(posctl, mode):
    """The validator's N_prev instance as one fused solve with the full D side (Dmin = 1: all C(64,7) =
    621 216 192 records) and the unpruned 30-prime I side: the partitioned mode buckets the D side as
    several range tables (each within the window build's reach) and probes every I record against each;
    the global-atomic mode takes one table. Both must give the composition's pair set."""
    c = I.control_instance("N_prev", nins=30)
    base = c.inst
    inst = I.Instance(M=base.M, r=base.r, B0=base.B0, INS=base.INS, P0=base.P0, Icap=prod(base.INS[-base.r:]) + 1,
                      Dmin=1, split=None, hmax=base.r, tag="N_prev_full_D")
    fx = posctl["compose"]["N_prev"]
    say(f"full D side in range tables ({mode}): M={inst.M} r={inst.r} nD={fx['nD']} nI={fx['nI']}")
    res = pipeline.solve(inst, mode=mode)
    assert (res.nD, res.nI) == (fx["nD"], fx["nI"]) == (comb(64, 7), comb(30, 7))
    d_dec, i_dec = pipeline.decoders(inst)
    pairs = sorted([sorted(inst.B0[x] for x in d_dec(int(a))), sorted(inst.INS[x] for x in i_dec(int(b)))]
                   for a, b in zip(res.pairs_d, res.pairs_i))
    assert pairs == fx["pairs"]
