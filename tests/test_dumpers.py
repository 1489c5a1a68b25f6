"""The record dumpers' wire format (cfjoin/dumpers.py): stdin, exit statuses, stderr lines and
stdout bytes against the reference dumpers compiled from their sources (oracle/_ref:
src/mitm64_ddump.cpp, src/mitm64_idump_stream2.cpp) and against the SHA-256 of the reference
streams in tests/golden/ref_streams.json. Rejected inputs and inputs with no records never reach
the device, so those cases run on the CPU; the record streams are GPU tests."""
import hashlib
import io
import json
import os
import subprocess
from math import prod

import pytest

import oracle as O
from cfjoin import dumpers as DU
from cfjoin import instance as I

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
M0 = 1768248177696000


def ref_run(prog, text, env):
    path = os.path.join(O.REF_DIR, prog)
    if not os.path.exists(path):
        pytest.skip(f"{prog} is not built (oracle/Makefile, target ref)")
    pr = subprocess.run([path], input=text.encode(), capture_output=True,
                        env=dict(os.environ, **env), timeout=120)
    return pr.returncode, pr.stdout, pr.stderr.decode()


def ours(run, text, env):
    out, err = io.BytesIO(), io.StringIO()
    st = run(text, out, err, env)
    return st, out.getvalue(), err.getvalue()


def d_text(M, r, Dmin, B):
    return f"{M} {r} {Dmin}\n{len(B)} " + " ".join(map(str, B)) + "\n"


def i_text(M, r, P0modM, Icap, INS, T=1, B=()):
    return "\n".join([f"{M} {r} {P0modM} {Icap} 1 {T} 0", f"{len(B)} " + " ".join(map(str, B)),
                      f"{len(INS)} " + " ".join(map(str, INS))]) + "\n"


def base():
    return I.pool(M0, 2000)[:64]


def ins(n):
    """n primes of M0's pool past its base (the pool holds a few hundred)."""
    P = I.pool(M0, 40000)
    assert len(P) >= 64 + n
    return P[64:64 + n]


def ascending(n):
    """n ascending odd numbers: the reference checks counts and order before it reads a value."""
    return list(range(3, 3 + 2 * n, 2))


# ---- CPU: inputs the reference rejects, and inputs with no records --------------------------
D_BAD = [
    ("63 base primes", lambda B: d_text(M0, 3, 1, B[:63]), {}),
    ("r = 0", lambda B: d_text(M0, 0, 1, B), {}),
    ("r = 17", lambda B: d_text(M0, 17, 1, B), {}),
    ("not ascending", lambda B: d_text(M0, 3, 1, B[:10] + [B[11], B[10]] + B[12:]), {}),
    ("truncated", lambda B: d_text(M0, 3, 1, B).replace(" ".join(map(str, B)), " ".join(map(str, B[:61]))), {}),
    ("empty", lambda B: "", {}),
]


@pytest.mark.parametrize("name,make,env", D_BAD, ids=[c[0] for c in D_BAD])
def test_ddump_rejects_like_reference(name, make, env):
    text = make(base())
    st, out, err = ref_run("mitm64_ddump", text, env)
    assert st != 0
    assert ours(DU.run_ddump, text, env) == (st, out, err)


I_BAD = [
    ("300 numbers, packed ids", lambda: i_text(M0, 3, 5, 10 ** 12, ascending(300)), {"IDUMP_RANK": "0"}),
    ("r = 9, packed ids", lambda: i_text(M0, 9, 5, 10 ** 30, ascending(40)), {}),
    ("r = 17", lambda: i_text(M0, 17, 5, 10 ** 30, ascending(40)), {"IDUMP_RANK": "1"}),
    ("rank ids overflow", lambda: i_text(M0, 16, 5, 10 ** 30, ascending(1024)), {"IDUMP_RANK": "1"}),
    ("not ascending", lambda: i_text(M0, 3, 5, 10 ** 12, ascending(20)[::-1]), {}),
    ("truncated", lambda: i_text(M0, 3, 5, 10 ** 12, ascending(20)).rsplit(" ", 3)[0], {}),
]


@pytest.mark.parametrize("name,make,env", I_BAD, ids=[c[0] for c in I_BAD])
def test_idump_rejects_like_reference(name, make, env):
    text = make()
    st, out, err = ref_run("mitm64_idump_stream2", text, env)
    assert st != 0
    assert ours(DU.run_idump, text, env) == (st, out, err)


@pytest.mark.parametrize("Icap,r,T,env", [
    (1, 3, 1, {}),                                            # Icap <= 1: no records
    (0, 3, 4, {"IDUMP_SPLIT": "4000", "IDUMP_HMAX": "2"}),
    (10 ** 20, 12, 2, {"IDUMP_RANK": "1"}),                   # r beyond the pool: no records
])
def test_idump_empty_streams_like_reference(Icap, r, T, env):
    text = i_text(M0, r, 7, Icap, ins(20) if r <= 8 else ins(10), T=T)
    st, out, err = ref_run("mitm64_idump_stream2", text, env)
    assert st == 0 and out == b""
    assert ours(DU.run_idump, text, env) == (st, out, err)


def test_parse_matches_reference_fields():
    """The parsed request carries the values the reference reads (split index from the first prime
    above IDUMP_SPLIT, hmax from IDUMP_HMAX, Icap mod 2^128, ids forced to masks for r > 10)."""
    P = ins(30)
    q = DU.parse_idump(i_text(M0, 4, 9, (1 << 130) + 5, P, T=0), {"IDUMP_SPLIT": str(P[7]), "IDUMP_HMAX": "2"})
    assert (q.split_idx, q.hmax, q.Icap, q.T, q.rank) == (8, 2, 5, 1, False)
    d = DU.parse_ddump(d_text(M0, 11, 12345, base()), {"DDUMP_BITMASK": "0"})
    assert d.bitmask and d.r == 11 and d.Dmin == 12345


# ---- GPU: record streams -------------------------------------------------------------------
def _fixtures():
    with open(os.path.join(GOLDEN, "ref_streams.json")) as f:
        return [c for c in json.load(f)["cases"] if c["kind"] == "family"]


def _sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.gpu
@pytest.mark.parametrize("c", _fixtures(), ids=lambda c: f"{c['family']}-{c['D']}-r{c['r']}")
def test_streams_equal_reference_dumpers(workloads, c):
    f = workloads["families"][c["family"]]
    row = next(x for x in f["rows"] if int(x["D"]) == c["D"])
    inst = I.family_instance(c["D"], int(row["cap"]), int(f["U"]), c["r"])
    print("CASE dumpers", c["family"], c["D"], c["r"], flush=True)
    st, out, err = ours(DU.run_ddump, d_text(inst.M, c["r"], inst.Dmin, inst.B0), {"DDUMP_BITMASK": "0"})
    assert st == 0 and len(out) == 16 * c["nD"] and _sha(out) == c["d_sha"]
    assert err == f"TOTAL {c['nD']} records (r={c['r']}, filtered by prodD >= Dmin, ids=6-bit)\n"
    env = {"IDUMP_SPLIT": str(inst.split), "IDUMP_HMAX": str(inst.hmax), "IDUMP_RANK": "1" if c["rank"] else "0"}
    st, out, err = ours(DU.run_idump, i_text(inst.M, c["r"], inst.P0 % inst.M, inst.Icap, inst.INS), env)
    assert st == 0 and len(out) == 16 * c["nI"] and _sha(out) == c["i_sha"]
    assert err.splitlines()[-1].startswith(f"TOTAL {c['nI']} records (r={c['r']}, |INS|={len(inst.INS)}, ")


def threads_case():
    """The threaded I-dump input of the test below (four threads, a stratum, 8-bit ids)."""
    P = ins(120)
    text = i_text(M0, 4, 123456789 % M0, prod(P[:4]) * 40, P, T=4)
    return text, {"IDUMP_SPLIT": str(P[60]), "IDUMP_HMAX": "2", "IDUMP_RANK": "0"}


def mask_cases():
    """D dumps with 64-bit mask ids: by DDUMP_BITMASK at r = 5, and forced by r = 11."""
    B = base()
    return [(5, d_text(M0, 5, prod(B[-5:]) // 50, B), {"DDUMP_BITMASK": "1"}),
            (11, d_text(M0, 11, prod(B[-11:]) // 4, B), {})]


@pytest.mark.gpu
def test_threaded_counts_and_masks_like_reference():
    """T > 1: the per-thread count lines and TOTAL equal the reference's (its records interleave by
    thread, so the streams are compared as record sets); D ids as 64-bit masks, by env and for r > 10.
    The reference programs' output is held in tests/golden/dumper_threads.json (written from the
    compiled programs by tests/golden/make_dumper_threads.py), so the test needs no build of them."""
    import numpy as np
    with open(os.path.join(GOLDEN, "dumper_threads.json")) as f:
        fx = json.load(f)
    text, env = threads_case()
    print("CASE dumpers threads", flush=True)
    st, out, err = ours(DU.run_idump, text, env)
    exp = fx["idump_threads"]
    assert (st, err) == (exp["status"], exp["stderr"])
    b = np.frombuffer(out, np.uint64).reshape(-1, 2)
    b = np.ascontiguousarray(b[np.lexsort((b[:, 1], b[:, 0]))])
    assert b.shape[0] == exp["records"] > 0 and _sha(b.tobytes()) == exp["sorted_sha"]
    for (r, text, env), exp in zip(mask_cases(), fx["ddump_masks"]):
        print("CASE dumpers mask", r, flush=True)
        st, out, err = ours(DU.run_ddump, text, env)
        assert exp["r"] == r and exp["bytes"] > 0
        assert (st, err, len(out), _sha(out)) == (exp["status"], exp["stderr"], exp["bytes"], exp["sha"])
