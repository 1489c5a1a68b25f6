"""CPU tests of the oracle (oracle/oracle.c): how each restated routine is pinned.

  or_ddump / or_idump  byte-identical record streams of the compiled reference dumpers
                       (tests/golden/ref_streams.json, small_records.npz), the logged counts of the
                       reference's own runs (tests/golden/logged_instances.json), brute force over
                       itertools.combinations on small pools
  or_gen_rank          the lexicographic order of itertools.combinations (rank order) and the
                       closed form mult * prod mod M
  or_join_sorted       nested-loop brute force and the numpy searchsorted restatement
  or_cpu_join          logged pair counts, and the MATCH sets of the compiled reference CPU join on
                       the two r = 6 anchors (every reference match is in the full pair set)
Several tests plant a mistake in an oracle input or output and check that the pin catches it.
"""
import hashlib
import itertools
import json
import os
from math import prod

import numpy as np
import pytest

import oracle as O
from cfjoin import instance as I
from cfjoin import verify as V
from cfjoin import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def stream(k, i):
    return np.stack([k, i], axis=1).astype(np.uint64)


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(GOLDEN, "ref_streams.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def logged():
    return W.logged_table(os.path.join(GOLDEN, "logged_instances.json"))


def fam_inst(workloads, fam, D, r):
    f = workloads["families"][fam]
    row = next(x for x in f["rows"] if int(x["D"]) == D)
    return I.family_instance(D, int(row["cap"]), int(f["U"]), r)


# ---- streams of the compiled reference dumpers --------------------------------------------
def test_family_streams_match_reference_dumpers(fixtures, workloads):
    for c in fixtures["cases"]:
        if c["kind"] != "family":
            continue
        inst = fam_inst(workloads, c["family"], c["D"], c["r"])
        dk, di = O.ddump(inst.M, inst.r, inst.Dmin, inst.B0)
        ik, ii = O.idump(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax,
                         rank=c["rank"])
        assert (len(dk), len(ik)) == (c["nD"], c["nI"])
        assert sha(stream(dk, di)) == c["d_sha"], c
        assert sha(stream(ik, ii)) == c["i_sha"], c


def test_c1_streams_match_reference_dumper(fixtures):
    for c in fixtures["cases"]:
        if c["kind"] != "c1":
            continue
        inst = I.c1_instance(c["seed"], c["s"])
        ik, ii = O.idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, None, None, rank=False)
        assert len(ik) == c["nI"] and sha(stream(ik, ii)) == c["i_sha"]


def test_small_record_lists(workloads):
    z = np.load(os.path.join(GOLDEN, "small_records.npz"))
    inst = I.c1_instance(0, 10)
    ik, ii = O.idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, None, None, rank=False)
    assert np.array_equal(stream(ik, ii), z["c1_seed0_s10_I"])
    inst = fam_inst(workloads, "M0", 1014485472000, 5)
    dk, di = O.ddump(inst.M, 5, inst.Dmin, inst.B0)
    assert np.array_equal(stream(dk, di), z["M0fam_D1014485472000_r5_D"])


def test_planted_key_error_is_caught(fixtures, workloads):
    c = next(c for c in fixtures["cases"] if c["kind"] == "family")
    inst = fam_inst(workloads, c["family"], c["D"], c["r"])
    dk, di = O.ddump(inst.M, inst.r, inst.Dmin, inst.B0)
    dk = dk.copy()
    dk[len(dk) // 2] ^= 1
    assert sha(stream(dk, di)) != c["d_sha"]


# ---- logged counts --------------------------------------------------------------------------
def _small_logged(logged_rows, limit_keys=3_000_000):
    rows = [r for r in logged_rows.values() if r["path"] in ("dfilter", "host-exact") and r["nD"] + r["nI"] <= limit_keys]
    rows.sort(key=lambda r: (r["family"], r["D"], r["r"]))
    return rows[::max(1, len(rows) // 40)]


def test_logged_counts_reproduced(logged, workloads):
    for r in _small_logged(logged):
        inst = fam_inst(workloads, r["family"], r["D"], r["r"])
        assert inst is not None, r
        assert len(inst.INS) == r["nINS"] and inst.hmax == r["hmax"], r
        res = O.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                         inst.hmax, inst.rank_ids, threads=4)
        assert (res["nD"], res["nI"], len(res["pairs_d"])) == (r["nD"], r["nI"], r["pairs"]), r


# ---- brute force on small pools ----------------------------------------------------------
SMALL_PRIMES = [p for p in range(3, 400) if all(p % q for q in range(2, int(p ** 0.5) + 1))]


def brute_d(B, r, M, Dmin):
    """D records in the dumper's order: descending index choices (src/mitm64_ddump.cpp:36-41)."""
    out = []
    for c in itertools.combinations(range(len(B) - 1, -1, -1), r):
        if prod(B[x] for x in c) >= Dmin:
            idx = sorted(c)
            out.append((prod(B[x] for x in c) % M, sum(x << (6 * p) for p, x in enumerate(idx))))
    return out


def brute_i(INS, r, M, P0modM, Icap, split_idx, hmax):
    """I records in ascending lexicographic order with colex ids."""
    from math import comb
    out = []
    for c in itertools.combinations(range(len(INS)), r):
        if prod(INS[x] for x in c) >= Icap:
            continue
        if sum(1 for x in c if x >= split_idx) > hmax:
            continue
        out.append((P0modM * prod(INS[x] for x in c) % M, sum(comb(x, j + 1) for j, x in enumerate(c))))
    return out


@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_ddump_brute_force(r):
    B = SMALL_PRIMES[:14]
    M = 2 ** 5 * 3 ** 2 * 5 * 7 * 11
    for Dmin in (1, prod(B[-r:]) // 7, prod(B[-r:]) // 2, prod(B[-r:]), prod(B[-r:]) + 1):
        dk, di = O.ddump(M, r, Dmin, B)
        exp = brute_d(B, r, M, Dmin)
        assert sorted(zip(dk.tolist(), di.tolist())) == sorted(exp)


@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_idump_brute_force_and_icap_exclusive(r):
    INS = SMALL_PRIMES[5:22]
    M = 2 ** 6 * 3 ** 3 * 5 ** 2 * 7 * 13
    P0m = 123457 % M
    for split_idx, hmax in ((len(INS), r), (9, 1), (4, 0), (12, 2)):
        # Icap equal to a product that occurs: that subset is excluded (exclusive bound)
        Icap = prod(INS[3:3 + r])
        k, i = O.idump(M, r, P0m, Icap, INS, split_idx, hmax, rank=True)
        exp = brute_i(INS, r, M, P0m, Icap, split_idx, hmax)
        assert list(zip(k.tolist(), i.tolist())) == exp
    # a planted inclusive bound is caught by the brute force
    k_incl, _ = O.idump(M, r, P0m, prod(INS[3:3 + r]) + 1, INS, len(INS), r, rank=True)
    assert len(k_incl) == len(brute_i(INS, r, M, P0m, prod(INS[3:3 + r]), len(INS), r)) + 1


def test_gen_rank_is_lexicographic_rank_order():
    primes = SMALL_PRIMES[:12]
    for r, M, mult, bits in ((3, 1768248177696000, 1, 6), (4, 10007, 5, 8), (2, (1 << 64) - 59, 3, 10)):
        k, i = O.gen_rank(primes, r, M, mult, bits)
        exp = [(mult * prod(primes[x] for x in c) % M, sum(x << (bits * p) for p, x in enumerate(c)))
               for c in itertools.combinations(range(len(primes)), r)]
        assert list(zip(k.tolist(), i.tolist())) == exp
        k2, _ = O.gen_rank(primes, r, M, mult, bits, lo=5, count=7)
        assert k2.tolist() == [e[0] for e in exp[5:12]]


def test_join_sorted_brute_force():
    rng = np.random.default_rng(3)
    for nD, nI, M in ((50, 80, 7), (200, 300, 97), (0, 5, 3), (5, 0, 3)):
        Dk = np.sort(rng.integers(0, M, nD).astype(np.uint64))
        Ik = np.sort(rng.integers(0, M, nI).astype(np.uint64))
        Did = np.arange(nD, dtype=np.uint64) * 3
        Iid = np.arange(nI, dtype=np.uint64) * 5 + 1
        od, oi = O.join_sorted(Dk, Did, Ik, Iid)
        exp = [(int(Did[d]), int(Iid[i])) for i in range(nI) for d in range(nD) if Dk[d] == Ik[i]]
        assert list(zip(od.tolist(), oi.tolist())) == exp
        lo = np.searchsorted(Dk, Ik, "left")
        hi = np.searchsorted(Dk, Ik, "right")
        assert len(od) == int((hi - lo).sum())


def test_cpu_join_equals_sorted_join_on_c1():
    for seed, s in ((0, 100), (2, None), (1, 10000)):
        inst = I.c1_instance(seed, s)
        Dk, Did, Ik, Iid = O.instance_records(inst)
        o = np.argsort(Dk, kind="stable")
        q = np.argsort(Ik, kind="stable")
        ed, ei = O.join_sorted(Dk[o], Did[o], Ik[q], Iid[q])
        res = O.cpu_join(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx, inst.hmax,
                         inst.rank_ids, threads=3)
        assert np.array_equal(O.pair_set(res["pairs_d"], res["pairs_i"]), O.pair_set(ed, ei))


@pytest.mark.slow
@pytest.mark.parametrize("fam", ["M0", "lambdaN"])
def test_anchor_pairs_contain_reference_matches(fam, fixtures, logged):
    """The reference CPU join keeps one D per residue (the largest product, src/mitm64.cpp:56), so
    its MATCH set must lie inside the full pair set, which must have the logged size."""
    inst = W.anchor(fam, 6, logged)
    res = O.cpu_join(inst.M, 6, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx, inst.hmax,
                     inst.rank_ids, threads=8)
    lg = inst.meta["logged"]
    assert (res["nD"], res["nI"], len(res["pairs_d"])) == (lg["nD"], lg["nI"], lg["pairs"])
    mine = set()
    for d, i in zip(res["pairs_d"], res["pairs_i"]):
        D = tuple(sorted(inst.B0[x] for x in V.decode_packed(int(d), 6, 6)))
        Iset = tuple(sorted(inst.INS[x] for x in V.unrank_colex(int(i), 6, len(inst.INS))))
        mine.add((D, Iset))
    anc = next(a for a in fixtures["anchors"] if a["family"] == fam)
    ref = {(tuple(a), tuple(b)) for a, b in anc["matches"]}
    assert ref and ref <= mine


def test_dsweep_join_equals_cpu_join(logged, workloads):
    """The swapped-role join (hash I, sweep D on several threads) returns the pair multiset, the
    record counts and hence the logged counts of the D-hashing join."""
    for r in _small_logged(logged)[::4] + [None]:
        if r is None:
            inst = I.c1_instance(0, 100)
        else:
            inst = fam_inst(workloads, r["family"], r["D"], r["r"])
        args = (inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx, inst.hmax,
                inst.rank_ids)
        a = O.cpu_join(*args, threads=2)
        b = O.join_dsweep(*args, threads=3)
        assert (a["nD"], a["nI"]) == (b["nD"], b["nI"])
        assert np.array_equal(O.pair_set(a["pairs_d"], a["pairs_i"]), O.pair_set(b["pairs_d"], b["pairs_i"]))
        if r is not None:
            assert (b["nD"], b["nI"], len(b["pairs_d"])) == (r["nD"], r["nI"], r["pairs"])


def _fixture_pairs_hold(inst, pairs):
    """Every fixture pair decodes to r distinct D and I indices whose residues agree mod M
    (P0 * prod(I) == prod(D)), with prod(D) >= Dmin and prod(I) < Icap, and the pair list has no repeats."""
    from cfjoin import pipeline
    d_dec, i_dec = pipeline.decoders(inst)
    assert len({tuple(p) for p in pairs}) == len(pairs)
    for dv, iv in pairs:
        d, i = d_dec(int(dv)), i_dec(int(iv))
        assert len(set(d)) == len(set(i)) == inst.r
        pd = prod(inst.B0[x] for x in d)
        pi = prod(inst.INS[x] for x in i)
        assert (inst.P0 * pi - pd) % inst.M == 0
        assert pd >= inst.Dmin and pi < inst.Icap


def test_large_radius_fixture_pairs_are_exact(logged):
    """The 78 pairs the swapped-role join found on the full-D r = 8 instance are real joins."""
    import copy
    with open(os.path.join(GOLDEN, "c4_full_pairs.json")) as f:
        fix = json.load(f)
    inst = copy.deepcopy(W.anchor("lambdaN", 8, logged))
    inst.Dmin = 1
    assert (fix["M"], fix["r"]) == (inst.M, 8) and len(fix["pairs"]) == 78
    _fixture_pairs_hold(inst, fix["pairs"])


def test_seeded_modulus_fixture_pairs_are_exact():
    """The pairs the CPU join found on the seeded k = 64 moduli (bound s = 1000) are real joins, on
    the moduli the generator still draws."""
    with open(os.path.join(GOLDEN, "c2_pairs.json")) as f:
        fix = json.load(f)
    for f in fix:
        M, cap = I.c2_modulus(f["seed"])
        assert M == f["M"]
        assert 200 <= len(I.pool(M, cap)) <= 500
        inst = I.c2_instance(f["seed"], f["s"])
        _fixture_pairs_hold(inst, f["pairs"])


# ---- the exact join in Python integers (moduli of any size) --------------------------------------
def _brute_join(M, r, P0, Icap, Dmin, B, INS, split_idx, hmax):
    Ds = [c for c in itertools.combinations(range(len(B)), r) if prod(B[i] for i in c) >= Dmin]
    Is = [c for c in itertools.combinations(range(len(INS)), r)
          if prod(INS[i] for i in c) < Icap and sum(i >= split_idx for i in c) <= hmax]
    pairs = {(d, i) for d in Ds for i in Is if (prod(B[x] for x in d) - P0 * prod(INS[x] for x in i)) % M == 0}
    return len(Ds), len(Is), pairs


@pytest.mark.parametrize("M", [(1 << 64) + 13 * 2 ** 40, 3 << 70, 2 ** 20 * 3 ** 5 * 5 ** 3 * 7 * 11, 997])
def test_exact_join_brute_force(M):
    """exact_join against all r-subsets: the bounds (Dmin inclusive, Icap exclusive, stratum) and the
    congruence, with a planted pair; a planted off-by-one in either bound would change the counts."""
    P = [p for p in range(5, 400) if V.is_prime(p) and M % p]
    B, INS = P[:12], P[12:30]
    r = 3
    D0, I0 = (9, 10, 11), (0, 4, 7)
    P0 = prod(B[i] for i in D0) * pow(prod(INS[i] for i in I0), -1, M) % M
    Dmin = prod(B[i] for i in D0)          # D0 sits exactly on the inclusive bound
    Icap = prod(INS[i] for i in (3, 15, 17))
    for split_idx, hmax in ((len(INS), r), (10, 1)):
        got = O.exact_join(M, r, P0, Icap, Dmin, B, INS, split_idx, hmax)
        want = _brute_join(M, r, P0, Icap, Dmin, B, INS, split_idx, hmax)
        assert got == want and (D0, I0) in got[2]
        assert O.exact_join(M, r, P0, Icap, Dmin + 1, B, INS, split_idx, hmax)[0] == want[0] - 1


def test_exact_join_equals_cpu_join_below_2_64():
    """Below 2^64 the exact join and the C oracle's hash join give the same records and pair set."""
    for seed, s in ((0, 100), (1, 10 ** 4), (3, 10)):
        inst = I.c1_instance(seed, s)
        nD, nI, pairs = O.exact_join(inst.M, inst.r, inst.P0, inst.Icap, inst.Dmin, inst.B0, inst.INS,
                                     inst.split_idx, inst.hmax)
        res = O.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                         inst.hmax, inst.rank_ids, threads=2)
        dd = lambda v: tuple(sorted(V.decode_packed(v, inst.r, 6)))  # noqa: E731
        di = (lambda v: V.unrank_colex(v, inst.r, len(inst.INS))) if inst.rank_ids else \
            (lambda v: V.decode_packed(v, inst.r, 8))
        theirs = {(dd(int(a)), tuple(di(int(b)))) for a, b in zip(res["pairs_d"], res["pairs_i"])}
        assert (nD, nI) == (res["nD"], res["nI"]) and len(theirs) == len(res["pairs_d"])
        assert pairs == theirs and len(pairs) > 0
