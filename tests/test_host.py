"""CPU tests of the host side: instance recipe, Korselt re-check, id codecs, drop-in preconditions,
the C ABI surface (the library loads and exports every declared symbol; no device calls)."""
import ctypes
import json
import os
import re
import subprocess
import sys
from math import comb, prod

import pytest

from cfjoin import instance as I
from cfjoin import lib as L
from cfjoin import verify as V
from cfjoin import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# ---- instance recipe against every logged instance ----------------------------------------
def test_recipe_reproduces_every_logged_pool_and_stratum(workloads):
    rows = json.load(open(os.path.join(GOLDEN, "logged_instances.json")))
    pools = {}
    checked = 0
    for r in rows:
        if r["path"] == "nothing":
            continue
        fam = workloads["families"][r["family"]]
        row = next(x for x in fam["rows"] if int(x["D"]) == r["D"])
        key = (r["family"], r["D"])
        if key not in pools:
            pools[key] = I.pool(r["D"], int(row["cap"]))
            assert len(pools[key]) == int(row["pool_capped"]), key
        inst = I.family_instance(r["D"], int(row["cap"]), int(fam["U"]), r["r"], P=pools[key])
        assert inst is not None, r
        assert len(inst.INS) == r["nINS"] and inst.hmax == r["hmax"], (r, len(inst.INS), inst.hmax)
        checked += 1
    assert checked == 1412


def test_audit_batches_cover_every_logged_instance_once():
    """The coverage audit's grouping (cfjoin/audit.py): every logged instance in exactly one batch,
    each batch within its key and instance limits (a single larger instance alone) and with the sum
    of its moduli below 2^64 - 1 (one key space per batch)."""
    from cfjoin import audit
    items = audit.logged_instances(os.path.join(GOLDEN, "logged_instances.json"))
    assert len(items) == 1412
    todo = [(r, x) for r, x in items if x.Icap.bit_length() <= audit.DEVICE_ICAP_BITS]
    seen = 0
    for g in audit._batches(todo, 2e9, 256):
        keys = sum(r["nD"] + r["nI"] for r, _ in g)
        assert len(g) == 1 or (keys <= 2e9 and len(g) <= 256)
        assert sum(x.M for _, x in g) < (1 << 64) - 1
        seen += len(g)
    assert seen == len(todo) == 1411


def test_batch_workload_shape(workloads):
    logged = W.logged_table(os.path.join(GOLDEN, "logged_instances.json"))
    insts = W.batch_instances("lambdaN", logged=logged)
    assert len(insts) == 280
    assert all(x.meta["logged"] is not None for x in insts)
    assert sum(x.meta["logged"]["nD"] + x.meta["logged"]["nI"] for x in insts) == 1601806908


def test_anchor_bounds():
    inst = I.anchor_instance("M0", 6)
    assert (len(inst.INS), inst.n_ins_full, inst.hmax, inst.split_idx) == (375, 461, 2, 107)
    # Icap is exclusive and at most topI + 1; Dmin follows P0 * Imin / U
    assert inst.Icap <= prod(inst.INS[-6:]) + 1 or inst.Icap <= prod(I.pool(inst.M, 82177)[64:][-6:]) + 1
    assert inst.Dmin == inst.P0 * prod(inst.INS[:6]) // inst.U + 1


def test_c1_counts_from_the_survey():
    got = []
    for s in (10, 100, 10 ** 4):
        inst = I.c1_instance(0, s)
        import oracle as O
        Dk, _, Ik, _ = O.instance_records(inst)
        got.append((len(Dk), len(Ik)))
    assert got == [(343, 1741), (2292, 17069), (3654, 80703)]


# ---- Korselt re-check ------------------------------------------------------------------------
def test_certificate_known_answers():
    assert V.verify_certificate([3, 11, 17])[:2] == (True, 561)
    assert V.verify_certificate([7, 13, 19])[:2] == (True, 1729)
    assert V.verify_certificate([7, 11, 13, 41])[:2] == (True, 41041)
    assert V.verify_certificate([5, 7, 17, 19, 73])[:2] == (True, 825265)
    assert not V.verify_certificate([3, 11, 16])[0]      # 16 is not prime
    assert not V.verify_certificate([3, 3, 11])[0]       # repeated factor
    assert not V.verify_certificate([3, 11])[0]          # fewer than 3 factors
    assert not V.verify_certificate([3, 5, 7])[0]        # 105: Korselt fails


def test_incumbents_verify(workloads):
    for name, inc in workloads["incumbents"].items():
        ok, n, why = V.verify_certificate(inc["factors"])
        assert ok and n == int(inc["n"]) and len(inc["factors"]) == 64, (name, why)
        lam = 1
        from math import lcm
        for p in inc["factors"]:
            lam = lcm(lam, p - 1)
        # Korselt through the modulus: lambda(n) divides the search modulus M
        assert int(inc["modulus"]) % lam == 0, name


def test_is_prime_agrees_with_sieve():
    N = 20000
    sieve = bytearray([1]) * (N + 1)
    sieve[0] = sieve[1] = 0
    for i in range(2, int(N ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    assert all(V.is_prime(n) == bool(sieve[n]) for n in range(N + 1))
    assert V.is_prime(2 ** 61 - 1) and not V.is_prime(3215031751)  # a strong pseudoprime to 2, 3, 5, 7


def test_id_codecs_round_trip():
    n, r = 40, 5
    import itertools
    for c in itertools.islice(itertools.combinations(range(n), r), 0, 20000, 37):
        rank = sum(comb(x, j + 1) for j, x in enumerate(c))
        assert V.unrank_colex(rank, r, n) == c
        packed = sum(x << (8 * p) for p, x in enumerate(c))
        assert V.decode_packed(packed, r, 8) == c
        assert V.decode_bitmask(sum(1 << x for x in c)) == c
    with pytest.raises(ValueError):
        V.unrank_colex(comb(n, r) + 10 ** 9, r, n)


def test_completion_checks_residue_identity():
    B0 = [3, 5, 7, 11]
    INS = [13, 17, 19]
    M = 8
    P0 = prod(B0)
    # prod(D) = 3, P0 * prod(I) = 1155 * 13 = 15015 == 7 (mod 8) but 3 != 7: a false pair raises
    with pytest.raises(AssertionError):
        V.completion(B0, INS, (0,), (0,), P0, M)


# ---- drop-in modules: preconditions before any device work --------------------------------
def test_dropin_preconditions_raise_assertion():
    import igen_gpu as IG
    import mitm_gpu2 as G
    with pytest.raises(AssertionError):
        G.generate([3, 5, 7], 11, 100)
    with pytest.raises(AssertionError):
        G.generate(list(range(3, 200, 2)), 3, 100, bits=6)
    with pytest.raises(AssertionError):
        G.generate([3, 5, 7], 2, 100, lo=2, hi=1)
    with pytest.raises(AssertionError):
        list(IG.generate_chunks([5, 3, 7, 11], 3, 100, 7, 10 ** 6))      # not ascending
    with pytest.raises(AssertionError):
        list(IG.generate_chunks([3, 5, 7, 11], 9, 100, 7, 10 ** 6))      # r > 8
    with pytest.raises(AssertionError):
        list(IG.generate_chunks([3, 5, 7, 11], 3, 100, 7, 1 << 129))     # Icap beyond 128 bits
    gen = IG.generate_chunks([3, 5, 7, 11], 3, 100, 7, 1)                # Icap <= 1: nothing, total 0
    with pytest.raises(StopIteration) as st:
        next(gen)
    assert st.value.value == 0


def test_quotient_table_closed_form():
    import igen_gpu as IG
    INS = [p for p in range(100, 400) if V.is_prime(p)]
    Icap = prod(INS[:4]) * 1000
    r = 5
    H = IG.quotient_table(INS, r, Icap)
    n = len(INS)
    for t in range(r):
        for x in range(n):
            exp = (Icap - 1) // prod(INS[x:x + t + 1]) if x + t < n else 0
            assert H[t * n + x] == exp


# ---- the C ABI ------------------------------------------------------------------------------
def header_functions():
    text = open(os.path.join(ROOT, "include", "cfjoin.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(cf_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 30
    so = ctypes.CDLL(L.SO_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)
    assert L.load().cf_version() == 1


def test_error_status_maps_to_reference_exceptions():
    lib = L.load()
    out = ctypes.c_void_p()
    # a NULL context is a violated precondition: -1 -> AssertionError, with a message
    rc = lib.cf_alloc(None, 16, ctypes.byref(out))
    assert rc == -1
    with pytest.raises(AssertionError, match="NULL"):
        L.check(rc, "cf_alloc")


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr


# ---- device-side arithmetic and bucket geometry, compiled for the host ----------------------
@pytest.mark.parametrize("name", ["t_modarith", "t_keyhash"])
def test_native_checks(tmp_path, name):
    """t_modarith: csrc/modarith.h (plain Montgomery step, Garner's key, the 32-bit scaled leaf step)
    against 128-bit integer arithmetic. t_keyhash: csrc/keyhash.h -- every hash's bucket lies in its
    own hash range and partition for the geometries of 1 to 10^10 D records, and a planted
    partition-scaled map is caught. Each program exits 0 when every check holds."""
    exe = tmp_path / name
    src = os.path.join(ROOT, "tests", "native", name + ".cpp")
    inc = os.path.join(ROOT, "carmichael-frontier_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), src], check=True, timeout=300)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout
    assert "bad=0" in p.stdout
