"""GPU parity: every device entry point against the oracle (the CPU restatement), bit-exact.

Run on an MI355X:  python -u -m pytest tests -m gpu -x -v --timeout 120 --timeout-method thread
Each test prints the case it is about to run (instance, sizes) before touching the device.
"""
import ctypes
import itertools
import os
from math import comb, prod

import numpy as np
import pytest

import oracle as O
from cfjoin import instance as I
from cfjoin import lib as L
from cfjoin import pipeline

pytestmark = pytest.mark.gpu

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), 'golden')
GOLDEN_LOGGED = os.path.join(GOLDEN_DIR, 'logged_instances.json')


def say(*a):
    print("CASE", *a, flush=True)


def dev(a):
    from cfjoin import xp
    return xp.asarray(np.asarray(a, np.uint64))


def family(workloads, fam, D, r):
    f = workloads["families"][fam]
    row = next(x for x in f["rows"] if int(x["D"]) == D)
    return I.family_instance(D, int(row["cap"]), int(f["U"]), r)


# ---- compatibility generate(): rank ranges ------------------------------------------------
@pytest.mark.parametrize("M,r,bits,mult", [
    (1768248177696000, 3, 6, 1),
    (1768248177696000, 4, 6, 1),
    (155016423578016000, 2, 10, 12345678901),
    (17297280, 3, 8, 7),                   # small even modulus, heavy key collisions
    ((1 << 64) - 59, 3, 6, 3),             # odd modulus near 2^64
    (1 << 61, 2, 6, 5),                    # pure power of two
    (1, 2, 6, 1),                          # M = 1: every key is 0
])
def test_generate_rank_ranges(M, r, bits, mult):
    import mitm_gpu2 as G
    primes = I.pool(1768248177696000, 700)[:64]
    total = comb(64, r)
    say("generate", M, r, bits, mult, total)
    k, i = G.generate(primes, r, M, mult, bits=bits)
    ek, ei = O.gen_rank(primes, r, M, mult, bits)
    assert np.array_equal(k.get(), ek) and np.array_equal(i.get(), ei)
    # rank chunks partition the rank space exactly (gpu/validate_mitm_gpu2.py:62-68)
    b = [total * j // 5 for j in range(6)]
    ks = [G.generate(primes, r, M, mult, bits=bits, lo=x, hi=y)[0].get() for x, y in zip(b, b[1:])]
    assert np.array_equal(np.concatenate(ks), ek)


def test_generate_preconditions():
    import mitm_gpu2 as G
    with pytest.raises(AssertionError):
        G.generate([3, 5, 7], 11, 100)
    with pytest.raises(AssertionError):
        G.generate(list(range(3, 200, 2)), 3, 100, bits=6)
    with pytest.raises(AssertionError):
        G.generate([3, 5, 7], 2, 100, lo=2, hi=1)
    k, i = G.generate([3, 5, 7], 2, 100, lo=1, hi=1)
    assert k.size == 0 and i.size == 0


# ---- filtered sides: enumeration plans vs the dumper restatements --------------------------
def plan_records(plan):
    from cfjoin import xp
    n = plan.total()
    k = xp.empty(n)
    i = xp.empty(n)
    if n:
        assert plan.write(0, plan.n, k.ptr, i.ptr) == n
    return k.get(), i.get()


@pytest.mark.parametrize("fam,D,r", [
    ("M0", 1768248177696000, 3),
    ("M0", 1768248177696000, 4),
    ("lambdaN", 10334428238534400, 6),
    ("M0", 1014485472000, 5),
    ("lambdaN", 155016423578016000, 5),
])
def test_plan_streams_match_dumpers(workloads, fam, D, r):
    inst = family(workloads, fam, D, r)
    say("plan", fam, D, r, len(inst.INS), inst.hmax)
    ek, ei = O.idump(inst.M, r, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax, rank=True)
    p = L.Plan.insertion(inst.INS, r, inst.M, inst.P0 % inst.M, inst.Icap, inst.split_idx, inst.hmax, L.CF_ID_COLEX)
    k, i = plan_records(p)
    assert len(k) == len(ek)
    assert np.array_equal(k, ek) and np.array_equal(i, ei), "I stream differs from the dumper order"
    fc = p.first_counts()
    exp = [len(O.idump(inst.M, r, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax, True, x, x + 1)[0])
           for x in range(len(inst.INS))]
    assert list(fc) == exp
    dk, di = O.ddump(inst.M, r, inst.Dmin, inst.B0)
    pd = L.Plan.deletion(inst.B0, r, inst.M, inst.Dmin, L.CF_ID_PACK_DESC)
    k, i = plan_records(pd)
    assert np.array_equal(k, dk) and np.array_equal(i, di), "D stream differs from the dumper order"


@pytest.mark.parametrize("r", [1, 2, 3, 5, 8])
def test_plan_packed_ids_and_strata(r):
    M = 1768248177696000
    P = I.pool(M, 6000)
    B0, INS = P[:64], P[64:64 + 200]
    P0 = prod(B0)
    split = 4000
    split_idx = next((j for j, q in enumerate(INS) if q > split), len(INS))
    INS = INS[:137]
    muls = {1: (50, 10 ** 3, 10 ** 4), 2: (50, 10 ** 3, 10 ** 4), 3: (50, 10 ** 3, 10 ** 4), 5: (3, 10, 20), 8: (2, 5, 6)}[r]
    for hmax, mul in zip((0, 1, r), muls):
        Icap = prod(INS[:r]) * mul
        say("packed", r, hmax, Icap.bit_length())
        ek, ei = O.idump(M, r, P0 % M, Icap, INS, split_idx, hmax, rank=False)
        p = L.Plan.insertion(INS, r, M, P0 % M, Icap, split_idx, hmax, L.CF_ID_PACK_ASC)
        k, i = plan_records(p)
        assert np.array_equal(k, ek) and np.array_equal(i, ei)


def test_plan_deletion_bitmask_and_edges():
    M = 155016423578016000
    B0 = I.pool(M, 2000)[:64]
    P0 = prod(B0)
    for r, Dmin, mode, bm in [(11, prod(B0[-11:]) // 4, L.CF_ID_MASK, True),
                              (3, 1, L.CF_ID_PACK_DESC, False),
                              (4, prod(B0[-4:]) + 1, L.CF_ID_PACK_DESC, False)]:
        say("ddump", r, Dmin.bit_length(), bm)
        ek, ei = O.ddump(M, r, Dmin, B0, bitmask=bm)
        k, i = plan_records(L.Plan.deletion(B0, r, M, Dmin, mode))
        assert np.array_equal(k, ek) and np.array_equal(i, ei)
    # Icap <= 1: empty plan
    p = L.Plan.insertion(B0, 3, M, P0 % M, 1, 64, 3)
    assert p.total() == 0


# ---- sort and join ------------------------------------------------------------------------
# 3e6 keys take two MSD passes before the per-bucket sort; 1e5 keys below 7 fill buckets past the
# LDS capacity, so the LSD sort finishes them
@pytest.mark.parametrize("n,M", [(1, 97), (1000, 97), (100000, 1 << 40), (300001, 1768248177696000), (70000, (1 << 64) - 59),
                                 (3000000, 1768248177696000), (100000, 7), (5000, 1 << 52)])
def test_sort_side_stable(n, M):
    import mitm_gpu2 as G
    rng = np.random.default_rng(n)
    k = rng.integers(0, M, n, dtype=np.uint64, endpoint=False) if M < (1 << 63) else rng.integers(0, 1 << 63, n, dtype=np.uint64) * 2
    i = np.arange(n, dtype=np.uint64)
    say("sort", n, M)
    ks, is_ = G.sort_side(dev(k), dev(i))
    o = np.argsort(k, kind="stable")
    assert np.array_equal(ks.get(), k[o]) and np.array_equal(is_.get(), i[o])


def _bunched(kind, rng):
    if kind == "half_in_one_range":     # half the keys below 2^40 of a 51-bit range
        k = rng.integers(0, 1 << 51, 300000, dtype=np.uint64)
        k[::2] = rng.integers(0, 1 << 40, 150000, dtype=np.uint64)
        return k
    if kind == "values_repeated":       # 1000 values, each 100 times
        return np.repeat(rng.integers(0, 1 << 40, 1000, dtype=np.uint64), 100)[rng.permutation(100000)]
    if kind == "all_equal":
        return np.full(50000, 123456789, np.uint64)
    if kind == "two_values":
        return rng.integers(0, 2, 200000, dtype=np.uint64) * np.uint64(1 << 45)
    if kind == "top_of_range":           # keys up to 2^64 - 1, a fifth of them exactly that
        k = rng.integers(0, 1 << 63, 150000, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        k[::5] = np.uint64(0xFFFFFFFFFFFFFFFF)
        return k
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["half_in_one_range", "values_repeated", "all_equal", "two_values", "top_of_range"])
def test_sort_side_bunched_keys(kind):
    """Keys bunched on a few ranges or values: buckets past the LDS take more top bits, groups of
    many equal keys go to the LSD passes in LDS, and a bucket of equal keys past the LDS to the LSD
    sort; each result is the stable argsort."""
    import mitm_gpu2 as G
    rng = np.random.default_rng(11)
    k = _bunched(kind, rng)
    i = np.arange(len(k), dtype=np.uint64)
    say("sort bunched", kind, len(k))
    ks, is_ = G.sort_side(dev(k), dev(i))
    o = np.argsort(k, kind="stable")
    assert np.array_equal(ks.get(), k[o]) and np.array_equal(is_.get(), i[o])


@pytest.mark.parametrize("n,M", [(3000000, 1 << 51), (100000, 1 << 40), (100000, 7), (4097, 1 << 63)])
def test_sort_pairs_in_place(n, M):
    """cf_sort_pairs (in place) against the stable argsort: with one MSD pass the in-place sort
    goes through a scratch pair, with two it returns to the input arrays."""
    rng = np.random.default_rng(n + 1)
    k = rng.integers(0, M, n, dtype=np.uint64)
    i = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    say("sort in place", n, M)
    dk, di = dev(k), dev(i)
    L.check(L.load().cf_sort_pairs(dk.ctx.handle, ctypes.c_void_p(dk.ptr), ctypes.c_void_p(di.ptr), n), "sort")
    o = np.argsort(k, kind="stable")
    assert np.array_equal(dk.get(), k[o]) and np.array_equal(di.get(), i[o])


def test_join_sorted_matches_oracle():
    import mitm_gpu2 as G
    rng = np.random.default_rng(5)
    # the unsorted case goes through a per-lane search; the last puts keys at 2^64 - 1 on both sides
    for nD, nI, M, sort_i in [(5000, 20000, 3000, True), (100, 100, 50, True), (200000, 300000, 1 << 30, True),
                              (1, 1, 2, True), (20, 7, 9, True), (4000, 9001, 2500, False),
                              (3000, 70000, 1 << 64, True)]:
        say("join", nD, nI, M, sort_i)
        Dk = rng.integers(0, M, nD, dtype=np.uint64)
        Ik = rng.integers(0, M, nI, dtype=np.uint64)
        if M == 1 << 64:
            Dk[:3] = np.uint64(0xFFFFFFFFFFFFFFFF)
            Ik[::3] = np.uint64(0xFFFFFFFFFFFFFFFF)
        Dk = np.sort(Dk)
        if sort_i:
            Ik = np.sort(Ik)
        Did = rng.integers(0, 1 << 60, nD, dtype=np.uint64)
        Iid = rng.integers(0, 1 << 60, nI, dtype=np.uint64)
        pd, pi = G.join_sorted(dev(Dk), dev(Did), dev(Ik), dev(Iid))
        ed, ei = O.join_sorted(Dk, Did, Ik, Iid)
        assert np.array_equal(pd.get(), ed) and np.array_equal(pi.get(), ei)


# ---- igen_gpu.generate_chunks: chunk boundaries and totals --------------------------------
def test_generate_chunks_boundaries(workloads):
    import igen_gpu as IG
    inst = family(workloads, "M0", 1768248177696000, 3)
    say("chunks", len(inst.INS))
    ek, ei = O.idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax, rank=True)
    per = [len(O.idump(inst.M, 3, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax, True, x, x + 1)[0])
           for x in range(len(inst.INS))]
    n = len(inst.INS)
    for max_records, block_bytes in [(50_000, 512 << 20), (10 ** 9, 8 * n * n * 7), (1, 8 * n * n)]:
        gen = IG.generate_chunks(inst.INS, 3, inst.M, inst.P0, inst.Icap, 4000, inst.hmax, max_records, block_bytes)
        sizes, ks = [], []
        while True:
            try:
                k, i = next(gen)
            except StopIteration as st:
                total = st.value
                break
            sizes.append(k.size)
            ks.append(k.get())
        # expected cut points, computed from the per-first-index counts
        ib = max(1, min(n, block_bytes // (8 * n * n)))
        exp = []
        a = 0
        while a < n:
            b = min(n, a + ib)
            a2 = a
            while a2 < b:
                b2, acc = a2 + 1, per[a2]
                while b2 < b and acc + per[b2] <= max_records:
                    acc += per[b2]
                    b2 += 1
                if acc:
                    exp.append(acc)
                a2 = b2
            a = b
        assert sizes == exp and total == len(ek)
        assert np.array_equal(np.concatenate(ks), ek)
    k, i, cnt = IG.generate(inst.INS, 3, inst.M, inst.P0, inst.Icap, 4000, inst.hmax)
    assert cnt == len(ek) and np.array_equal(k.get(), ek) and np.array_equal(i.get(), ei)


# ---- whole instances ----------------------------------------------------------------------
@pytest.mark.parametrize("fam,D,r,max_chunk", [
    ("lambdaN", 10334428238534400, 6, 0),
    ("M0", 1014485472000, 5, 0),
    ("M0", 1768248177696000, 4, 1_000_000),
    ("M0", 1768248177696000, 5, 0),
])
def test_solve_matches_cpu_join(workloads, fam, D, r, max_chunk):
    inst = family(workloads, fam, D, r)
    say("solve", fam, D, r, len(inst.INS), max_chunk)
    res = pipeline.solve(inst, max_chunk=max_chunk, mode="sortmerge" if max_chunk else "partition")
    ref = O.cpu_join(inst.M, r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx, inst.hmax,
                     inst.rank_ids, threads=8)
    assert (res.nD, res.nI) == (ref["nD"], ref["nI"])
    assert np.array_equal(res.pair_array(), O.pair_set(ref["pairs_d"], ref["pairs_i"]))
    if max_chunk:
        res2 = pipeline.solve(inst, max_chunk=max_chunk // 7 + 1, mode="partition")
        assert np.array_equal(res2.pair_array(), res.pair_array())
    for rec in pipeline.recheck(inst, res):
        assert rec["D"] and rec["I"]


@pytest.mark.parametrize("seed,s", [(0, 10), (0, 100), (1, 10 ** 4), (2, None), (3, 1)])
def test_solve_c1(seed, s):
    inst = I.c1_instance(seed, s)
    say("c1", seed, s, inst.M)
    res = pipeline.solve(inst)
    Dk, Did, Ik, Iid = O.instance_records(inst)
    o = np.argsort(Dk, kind="stable")
    q = np.argsort(Ik, kind="stable")
    ed, ei = O.join_sorted(Dk[o], Did[o], Ik[q], Iid[q])
    assert (res.nD, res.nI) == (len(Dk), len(Ik))
    assert np.array_equal(res.pair_array(), O.pair_set(ed, ei))


# ---- fused and batched solves ---------------------------------------------------------------
def test_partition_equals_sortmerge_and_logged(workloads):
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    for fam, D, r in [("M0", 1768248177696000, 5), ("lambdaN", 155016423578016000, 5), ("M0", 1014485472000, 5)]:
        inst = family(workloads, fam, D, r)
        say("partition", fam, D, r)
        a = pipeline.solve(inst, mode="partition")
        b = pipeline.solve(inst, mode="sortmerge")
        c = pipeline.solve(inst, mode="partition_dglobal")
        lg = logged[(fam, D, r)]
        assert (a.nD, a.nI, len(a.pairs_d)) == (b.nD, b.nI, len(b.pairs_d)) == (lg["nD"], lg["nI"], lg["pairs"])
        assert np.array_equal(a.pair_array(), b.pair_array())
        assert np.array_equal(c.pair_array(), b.pair_array())


def test_batch_equals_alone(workloads):
    """Every instance's pair set is identical solved alone and solved inside a mixed batch."""
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    insts = W.batch_instances("lambdaN", radii=(3, 4, 5, 6), logged=logged, limit=60)
    insts += [I.c1_instance(0, 100), I.c1_instance(1, None)]
    say("batch", len(insts))
    res, st = pipeline.solve_batch(insts)
    assert len(res) == len(insts)
    for inst, rb in zip(insts, res):
        ra = pipeline.solve(inst)
        assert (rb.nD, rb.nI) == (ra.nD, ra.nI), inst.tag
        assert np.array_equal(rb.pair_array(), ra.pair_array()), inst.tag
        lg = inst.meta.get("logged")
        if lg:
            assert (rb.nD, rb.nI, len(rb.pairs_d)) == (lg["nD"], lg["nI"], lg["pairs"]), inst.tag


@pytest.mark.parametrize("M", [64, 1024, 2 ** 20 + 7])
def test_solve_heavy_collisions(M):
    """Tiny moduli: keys repeat thousands of times, so buckets overflow their 8 table slots (slot 7
    holds the overflow mark and every probe of such a bucket is a candidate) and the probe finds more
    candidates than its first allowance (it is repeated with room for all); the pair multiset must
    still equal the oracle join, with the D side bucketed either way."""
    P = [p for p in range(3, 2000) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    B0, INS = P[:20], P[20:80]
    inst = I.Instance(M=M, r=3, B0=B0, INS=INS, P0=prod(B0), Icap=prod(INS[-3:]) + 1, Dmin=1, split=None, hmax=3,
                      tag=f"collide{M}")
    say("collide", M)
    Dk, Did, Ik, Iid = O.instance_records(inst)
    o = np.argsort(Dk, kind="stable")
    q = np.argsort(Ik, kind="stable")
    ed, ei = O.join_sorted(Dk[o], Did[o], Ik[q], Iid[q])
    for mode in ("partition", "partition_dglobal"):
        res = pipeline.solve(inst, mode=mode)
        assert (res.nD, res.nI) == (len(Dk), len(Ik))
        assert np.array_equal(res.pair_array(), O.pair_set(ed, ei)), mode


def test_solve_primes_above_32_bits():
    """Pools with primes above 2^32 take the 64-bit leaf step (no 32-bit scaled residues); a mixed
    batch holds instances of both kinds, each in its own key range."""
    from cfjoin import verify as V
    big = [q for q in range((1 << 32) + 1, (1 << 32) + 4000, 2) if V.is_prime(q)][:12]
    P = [p for p in range(3, 400) if V.is_prime(p)]
    M = 2 ** 8 * 3 ** 3 * 5 * 7      # small, so residues collide and pairs exist
    B0, INS = P[:24], P[24:60] + big
    inst = I.Instance(M=M, r=3, B0=B0, INS=INS, P0=prod(B0), Icap=prod(INS[-3:]) + 1, Dmin=1, split=None, hmax=3,
                      tag="wide")
    small = I.c1_instance(0, 100)
    say("wide", M, len(INS))
    res = pipeline.solve(inst)
    Dk, Did, Ik, Iid = O.instance_records(inst)
    o = np.argsort(Dk, kind="stable")
    q = np.argsort(Ik, kind="stable")
    ed, ei = O.join_sorted(Dk[o], Did[o], Ik[q], Iid[q])
    assert (res.nD, res.nI) == (len(Dk), len(Ik))
    assert len(ed) > 0 and np.array_equal(res.pair_array(), O.pair_set(ed, ei))
    both, _ = pipeline.solve_batch([inst, small])
    assert np.array_equal(both[0].pair_array(), res.pair_array())
    assert np.array_equal(both[1].pair_array(), pipeline.solve(small).pair_array())


@pytest.mark.parametrize("a", [31, 32, 33, 40, 62])
def test_solve_two_adic_part_around_32_bits(a):
    """M = 3 * 2^a: the leaf kernels carry x2 mod 2^a in 32 bits when 2^a <= 2^32 and in 64 bits
    above; P0 is chosen so that one (D, I) pair must meet (P0 = prod D0 / prod I0 mod M), and both
    paths, eager and batched, must return the oracle's pair set."""
    P = [p for p in range(5, 2000) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    B0, INS = P[:20], P[20:80]
    M = 3 << a
    D0, I0 = B0[3:6], INS[10:13]
    P0 = prod(D0) * pow(prod(I0), -1, M) % M
    inst = I.Instance(M=M, r=3, B0=B0, INS=INS, P0=P0, Icap=prod(INS[-3:]) + 1, Dmin=1, split=None, hmax=3,
                      tag=f"twoadic{a}")
    say("two-adic", a)
    Dk, Did, Ik, Iid = O.instance_records(inst)
    o = np.argsort(Dk, kind="stable")
    q = np.argsort(Ik, kind="stable")
    ed, ei = O.join_sorted(Dk[o], Did[o], Ik[q], Iid[q])
    assert len(ed) >= 1
    res = pipeline.solve(inst)
    assert (res.nD, res.nI) == (len(Dk), len(Ik))
    assert np.array_equal(res.pair_array(), O.pair_set(ed, ei))
    both, _ = pipeline.solve_batch([inst, I.c1_instance(0, 100)])
    assert np.array_equal(both[0].pair_array(), res.pair_array())


def test_batch_graph_replays_match_eager(workloads):
    """The captured batch (one HIP graph launch per solve) returns, on every replay, the pair sets and
    record counts of the eager solve -- with eager work and synchronizations between the replays.
    The batch mixes ordinary instances with one whose keys repeat (its chunk overflows the capacity
    layout and its probe outgrows the first candidate allowance, so those paths are captured too)."""
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    P = [p for p in range(3, 2000) if all(p % q for q in range(2, int(p ** 0.5) + 1))]
    collide = I.Instance(M=1024, r=3, B0=P[:20], INS=P[20:80], P0=prod(P[:20]), Icap=prod(P[77:80]) + 1, Dmin=1,
                         split=None, hmax=3, tag="collide1024")
    insts = W.batch_instances("lambdaN", radii=(4, 5, 6), logged=logged, limit=30)
    insts += [I.c1_instance(0, 100), collide]
    say("graph", len(insts))
    eager, _ = pipeline.solve_batch(insts)
    other = I.c1_instance(1, None)
    for max_chunk in (0, 50_000):
        b = pipeline.Batch(insts)
        b.capture(max_chunk)
        for rep in range(3):
            res, st = b.solve(max_chunk)
            assert len(res) == len(insts)
            for inst, ra, rb in zip(insts, eager, res):
                assert (rb.nD, rb.nI) == (ra.nD, ra.nI), (max_chunk, rep, inst.tag)
                assert np.array_equal(rb.pair_array(), ra.pair_array()), (max_chunk, rep, inst.tag)
            pipeline.solve(other)            # eager work between replays
            L.default_context().synchronize()
        b.close()


def test_autotune_only_changes_speed(workloads):
    """Every tuning candidate solves a mixed batch to the same counts and pair sets (the tuner checks
    each one against the default setting and raises otherwise); the chosen setting is then captured."""
    from cfjoin import autotune
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    insts = W.batch_instances("lambdaN", radii=(4, 5), logged=logged, limit=20) + [I.c1_instance(2, None)]
    say("tune", len(insts))
    eager, _ = pipeline.solve_batch(insts)
    b = pipeline.Batch(insts)
    tr = autotune.tune(b, reps=1, candidates={"slice_kib": [512, 4096], "d_per_bucket": [3, 8],
                                              "seg_fill_pct": [50, 90], "probe_blocks": [64, 4096]})
    assert len(tr["trials"]) == 9
    b.capture()
    res, _ = b.solve()
    for ra, rb in zip(eager, res):
        assert (ra.nD, ra.nI) == (rb.nD, rb.nI) and np.array_equal(ra.pair_array(), rb.pair_array())
    b.close()


def test_autotune_graph_timing_only_changes_speed(workloads):
    """The tuner timing captured replays (graph=True): each candidate is captured on a copy of the
    batch and replayed, its result compared with the default setting's; the batch itself stays
    uncaptured and takes the chosen setting."""
    from cfjoin import autotune
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    insts = W.batch_instances("lambdaN", radii=(4, 5), logged=logged, limit=12)
    say("tune graph", len(insts))
    eager, _ = pipeline.solve_batch(insts)
    b = pipeline.Batch(insts)
    # one record per bucket doubles the window width (2048 buckets): at more than 8 filter bits per
    # bucket a window's filter words then exceed the window sort's LDS array and take its per-word path
    tr = autotune.tune(b, reps=1, graph=True, candidates={"slice_kib": [512], "d_per_bucket": [1],
                                                          "seg_fill_pct": [90]})
    assert len(tr["trials"]) == 4
    assert b.tuning == tr["best"]
    b.capture()
    res, _ = b.solve()
    for ra, rb in zip(eager, res):
        assert (ra.nD, ra.nI) == (rb.nD, rb.nI) and np.array_equal(ra.pair_array(), rb.pair_array())
    b.close()


def test_server_results_equal_alone(workloads):
    """Requests served together by the continuous-batching server (several workers, batches of
    whatever has arrived) get exactly the counts and pair sets they get solved alone."""
    from cfjoin import serve
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    insts = W.serving_requests(24, seed=3, logged=logged, max_keys=3e7) + [I.c1_instance(0, 100)]
    say("serve", len(insts))
    srv = serve.Server(workers=3, max_batch=8)
    for k, x in enumerate(insts):
        srv.submit(x, k)
    got = {}
    while len(got) < len(insts):
        r = srv.next(30000)
        assert r is not None, "server produced no result within 30 s"
        assert r.status == 0, r.error
        got[r.tag] = r
    srv.close()
    assert max(r.batch_size for r in got.values()) > 1          # requests were batched together
    for k, x in enumerate(insts):
        alone = pipeline.solve(x)
        r = got[k]
        assert (r.nD, r.nI) == (alone.nD, alone.nI), x.tag
        assert np.array_equal(r.pair_array(), alone.pair_array()), x.tag
        lg = x.meta.get("logged")
        if lg:
            assert (r.nD, r.nI, len(r.pairs_d)) == (lg["nD"], lg["nI"], lg["pairs"]), x.tag


def test_server_tags_are_unique_until_released():
    """A tag names one request from its submission until its result is released: a second
    submission under it is refused while the first is queued or held (its result's arrays stay
    valid), and the tag is free again once released."""
    from cfjoin import serve
    inst = I.c1_instance(0, 100)
    say("serve tags")
    srv = serve.Server(workers=1, max_batch=4)
    try:
        srv.submit(inst, 5)
        with pytest.raises(AssertionError, match="tag already in use"):
            srv.submit(inst, 5)
        r = srv.next(30000)              # copies the pairs and releases the tag
        assert r is not None and r.status == 0 and r.tag == 5
        srv.submit(inst, 5)
        r2 = srv.next(30000)
        assert r2 is not None and r2.tag == 5 and np.array_equal(r2.pair_array(), r.pair_array())
    finally:
        srv.close()


def test_r8_anchor_d_filtered_matches_logged(workloads):
    """The large-radius anchor (lambda(N), r = 8, 256-prime pool, h <= 1) with the reference's
    D filter: 18 426 621 D and 471 754 092 I records, no pairs, as logged; I side in 8 chunks."""
    from cfjoin import workloads as W
    logged = W.logged_table(GOLDEN_LOGGED)
    inst = W.anchor("lambdaN", 8, logged)
    say("r8", inst.M, len(inst.INS))
    res = pipeline.solve(inst, max_chunk=1 << 26)
    lg = inst.meta["logged"]
    assert (res.nD, res.nI, len(res.pairs_d)) == (lg["nD"], lg["nI"], lg["pairs"]) == (18426621, 471754092, 0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_c2_seeded_moduli(seed):
    """Seeded k = 64, r = 6 moduli (pools of 200 primes after the cap is raised): bound s = 100 against
    the CPU join here; s = 1000 (6e8-1.7e9 records) against the pair sets the CPU join wrote into
    tests/golden/c2_pairs.json."""
    import json
    small = I.c2_instance(seed, 100)
    say("c2", seed, small.M)
    res = pipeline.solve(small)
    ref = O.cpu_join(small.M, 6, small.P0 % small.M, small.Icap, small.Dmin, small.B0, small.INS, small.split_idx,
                     small.hmax, small.rank_ids, threads=8)
    assert (res.nD, res.nI) == (ref["nD"], ref["nI"])
    assert np.array_equal(res.pair_array(), O.pair_set(ref["pairs_d"], ref["pairs_i"]))
    fix = next(f for f in json.load(open(os.path.join(GOLDEN_DIR, "c2_pairs.json"))) if f["seed"] == seed)
    big = I.c2_instance(seed, 1000)
    assert big.M == fix["M"] and len(big.INS) == fix["nINS"]
    res = pipeline.solve(big)
    assert (res.nD, res.nI) == (fix["nD"], fix["nI"])
    exp = np.array(fix["pairs"], dtype=np.uint64).reshape(-1, 2)
    assert np.array_equal(res.pair_array(), exp)


def test_r8_full_d_side_matches_oracle_fixture():
    """The large-radius instance with the full D side (all C(64,8) = 4 426 165 368 D subsets, about
    134 GiB of device memory) against the pair set the CPU oracle's swapped-role join wrote into
    tests/golden/c4_full_pairs.json; I side in 2^26-record chunks."""
    import copy
    import json
    from cfjoin import workloads as W
    fix = json.load(open(os.path.join(GOLDEN_DIR, "c4_full_pairs.json")))
    logged = W.logged_table(GOLDEN_LOGGED)
    inst = copy.deepcopy(W.anchor("lambdaN", 8, logged))
    inst.Dmin = 1
    say("r8-full", inst.M)
    res = pipeline.solve(inst, max_chunk=1 << 26)
    assert (res.nD, res.nI) == (fix["nD"], fix["nI"]) == (4426165368, 471754092)
    assert np.array_equal(res.pair_array(), np.array(fix["pairs"], dtype=np.uint64).reshape(-1, 2))
    L.default_context().release_pool()


# ---- coverage audit: every logged instance ----------------------------------------------------
def test_every_logged_instance_matches_log():
    """All 1 412 logged lines of both families (gpu/mitm_gpu_mstar_driver.py:241,284): each instance
    rebuilt from the census and solved on the device in batches; (nD, nI, pairs) equal to the log and
    every pair re-checked exactly. The one instance whose insertion bound needs 134 bits (M0, r = 14)
    is beyond the device's 128-bit products and takes the host-exact join, as the reference's did."""
    from cfjoin import audit
    say("audit", "all logged instances")
    rep = audit.audit(GOLDEN_LOGGED)
    assert rep["instances"] == 1412 and rep["solved"] == 1411 and rep["solved_host_exact"] == 1
    assert rep["mismatches"] == [] and rep["recheck_failures"] == 0
    assert [(b["family"], b["r"], b["icap_bits"]) for b in rep["beyond_device"]] == [("M0", 14, 134)]
