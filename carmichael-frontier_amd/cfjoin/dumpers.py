"""The record dumpers' wire format, generated on the GPU.

The reference's drivers read the D and I record sets from two CPU programs over pipes
(gpu/mitm_gpu_mstar_driver.py:101-107, 157, 245-249):

  src/mitm64_ddump.cpp:57-68          stdin  "M r Dmin\\n nB B_0 .. B_63\\n"; env DDUMP_BITMASK
  src/mitm64_idump_stream2.cpp:72-99  stdin  "M r P0modM Icap <word> T maxm\\n nB B..\\n nI INS..\\n";
                                      env IDUMP_SPLIT, IDUMP_HMAX, IDUMP_RANK

Both write the records to stdout as (key, id) pairs of little-endian u64, and a TOTAL line to stderr.
This module reads the same input, applies the same checks with the same exit statuses and
messages, and writes the same bytes: the records come from a device plan (cf_plan_create_d /
cf_plan_create_i), which enumerates them in the dumpers' depth-first order, so the D stream and
the single-threaded I stream are byte for byte the reference's. With T > 1 the reference's threads
interleave their buffers in whatever order they finish; here the records keep the single-threaded
order and the per-thread counts on stderr are those of its striping (thread t takes the first
indices t, t + T, ...).

Run as `python -m cfjoin.dumpers ddump` or `python -m cfjoin.dumpers idump` (stdin, stdout and
stderr as the reference programs).
"""
from __future__ import annotations

import os
import sys
from dataclasses import dataclass
from typing import BinaryIO, List, Mapping, Optional, TextIO

import numpy as np

U128 = 1 << 128


class DumpError(Exception):
    """An input the reference program rejects: its exit status and stderr message (may be empty)."""

    def __init__(self, status: int, message: str = ""):
        super().__init__(message)
        self.status = status
        self.message = message


def _atoi(s: Optional[str]) -> int:
    """C atoi: optional sign and leading digits, 0 if none."""
    if s is None:
        return 0
    s = s.lstrip()
    sign, i = 1, 0
    if s[:1] in ("+", "-"):
        sign = -1 if s[0] == "-" else 1
        i = 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def _strtoull(s: str) -> int:
    """C strtoull(s, NULL, 10) for the values the drivers pass: leading digits, 0 if none."""
    s = s.lstrip()
    j = 0
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j else 0


def _parse_u128(tok: str) -> int:
    """The dumpers' parse_u128: decimal digits, value reduced mod 2^128 (src/mitm64_ddump.cpp:26)."""
    v = 0
    for ch in tok:
        v = (v * 10 + (ord(ch) - 48)) % U128
    return v


class _Tokens:
    """scanf-like reading of whitespace-separated tokens; a failed read is exit status 2."""

    def __init__(self, text: str):
        self.t = text.split()
        self.i = 0

    def word(self) -> str:
        if self.i >= len(self.t):
            raise DumpError(2)
        self.i += 1
        return self.t[self.i - 1][:159]

    def integer(self, bits: int = 64, signed: bool = False) -> int:
        w = self.word()
        k = 1 if w[:1] in ("+", "-") else 0
        if not w[k:k + 1].isdigit():
            raise DumpError(2)
        j = k
        while j < len(w) and w[j].isdigit():
            j += 1
        v = int(w[:j])
        return v if signed else v % (1 << bits)


@dataclass
class DRequest:
    M: int
    r: int
    Dmin: int
    B: List[int]
    bitmask: bool


@dataclass
class IRequest:
    M: int
    r: int
    P0modM: int
    Icap: int
    T: int
    INS: List[int]
    split_idx: int
    hmax: int
    rank: bool


def parse_ddump(text: str, env: Mapping[str, str]) -> DRequest:
    """src/mitm64_ddump.cpp:57-64: input, DDUMP_BITMASK (forced for r > 10), checks."""
    tk = _Tokens(text)
    M = tk.integer()
    r = tk.integer(signed=True)
    Dmin = _parse_u128(tk.word())
    nB = tk.integer(signed=True)
    B = [tk.integer() for _ in range(max(nB, 0))]
    bitmask = _atoi(env.get("DDUMP_BITMASK")) != 0 if "DDUMP_BITMASK" in env else False
    if r > 10:
        bitmask = True
    if nB != 64 or r < 1 or r > 16:
        raise DumpError(3, "need 64 base primes and 1<=r<=16\n")
    if any(B[i] <= B[i - 1] for i in range(1, nB)):
        raise DumpError(4, "base must be strictly ascending\n")
    if M < 1:
        raise DumpError(3, "need M >= 1\n")   # the reference divides by M; here refused up front
    return DRequest(M, r, Dmin, B, bitmask)


def _binom(n: int, k: int) -> int:
    if k < 0 or k > n:
        return 0
    c = 1
    for j in range(k):
        c = c * (n - j) // (j + 1)
    return c


def parse_idump(text: str, env: Mapping[str, str]) -> IRequest:
    """src/mitm64_idump_stream2.cpp:72-84: input, IDUMP_RANK / IDUMP_SPLIT / IDUMP_HMAX, checks."""
    tk = _Tokens(text)
    M = tk.integer()
    r = tk.integer(signed=True)
    P0modM = tk.integer()
    Icap = _parse_u128(tk.word())
    tk.word()                                  # the product of the r smallest base primes: unused
    T = tk.integer(signed=True)
    tk.integer(signed=True)                    # maxm: unused
    nB = tk.integer(signed=True)
    for _ in range(max(nB, 0)):
        tk.integer()
    nI = tk.integer(signed=True)
    INS = [tk.integer() for _ in range(max(nI, 0))]
    rank = _atoi(env.get("IDUMP_RANK")) != 0 if "IDUMP_RANK" in env else False
    if nI > 1024 or r > 16 or (not rank and (nI > 256 or r > 8)):
        raise DumpError(3, "8-bit packing needs nI<=256, r<=8; IDUMP_RANK=1 allows nI<=1024, r<=16 "
                           "(ids must satisfy C(nI,r) < 2^64)\n")
    if rank and nI >= 0 and r >= 0 and _binom(nI, r) > (1 << 64) - 1:
        raise DumpError(5, f"rank ids overflow u64: C({nI},{r}) >= 2^64 -- shrink the pool (per-radius cap)\n")
    if any(INS[i] <= INS[i - 1] for i in range(1, nI)):
        raise DumpError(4, "INS must be strictly ascending\n")
    if r < 1 or M < 1:
        # undefined in the reference (it recurses past depth r, or divides by M = 0): refused here.
        # r above the pool size is accepted and gives no records, as there.
        raise DumpError(3, "need r >= 1 and M >= 1\n")
    split_idx = 1 << 30
    if "IDUMP_SPLIT" in env:
        sp = _strtoull(env["IDUMP_SPLIT"])
        split_idx = next((i for i, q in enumerate(INS) if q > sp), nI)
    hmax = _atoi(env["IDUMP_HMAX"]) if "IDUMP_HMAX" in env else 1 << 30
    return IRequest(M, r, P0modM, Icap, max(T, 1), INS, split_idx, hmax, rank)


# ---- record streams -------------------------------------------------------------------------
CHUNK_RECORDS = 1 << 26   # records per device write and host copy (1 GiB of keys and ids)


def _ranges(counts: np.ndarray, order: range):
    """Groups of consecutive first indices (in stream order) holding at most CHUNK_RECORDS records,
    or one index that alone holds more."""
    lo = None
    acc = 0
    last = None
    for x in order:
        c = int(counts[x])
        if lo is not None and acc + c > CHUNK_RECORDS:
            yield lo, last, acc
            lo, acc = None, 0
        if lo is None:
            lo = x
        acc += c
        last = x
    if lo is not None:
        yield lo, last, acc


def _write_plan(plan, order: range, out: BinaryIO) -> int:
    from . import xp
    counts = plan.first_counts()
    total = 0
    for a, b, n in _ranges(counts, order):
        if n == 0:
            continue
        lo, hi = (a, b + 1) if a <= b else (b, a + 1)
        k = xp.empty(n)
        i = xp.empty(n)
        got = plan.write(lo, hi, k.ptr, i.ptr)
        if got != n:
            raise RuntimeError(f"plan wrote {got} records for first indices [{lo}, {hi}), expected {n}")
        rec = np.empty((n, 2), np.uint64)
        rec[:, 0] = k.get()
        rec[:, 1] = i.get()
        out.write(rec.tobytes())
        total += n
    return total


def run_ddump(text: str, out: BinaryIO, err: TextIO, env: Optional[Mapping[str, str]] = None) -> int:
    """The D dumper (src/mitm64_ddump.cpp): records of the r-subsets D of the base with
    prod D >= Dmin, key prod D mod M, id 6-bit packed ascending (or the 64-bit deletion mask);
    depth-first with indices descending. Returns the exit status."""
    from . import lib as L
    env = os.environ if env is None else env
    try:
        q = parse_ddump(text, env)
    except DumpError as e:
        err.write(e.message)
        return e.status
    plan = L.Plan.deletion(q.B, q.r, q.M, q.Dmin, L.CF_ID_MASK if q.bitmask else L.CF_ID_PACK_DESC)
    try:
        # the depth-0 index runs from the top of the base down
        total = _write_plan(plan, range(len(q.B) - 1, -1, -1), out)
    finally:
        plan.close()
    out.flush()
    err.write(f"TOTAL {total} records (r={q.r}, filtered by prodD >= Dmin, ids={'bitmask' if q.bitmask else '6-bit'})\n")
    return 0


def run_idump(text: str, out: BinaryIO, err: TextIO, env: Optional[Mapping[str, str]] = None) -> int:
    """The I dumper (src/mitm64_idump_stream2.cpp): records of the r-subsets I of INS with
    prod I < Icap and at most hmax primes past the split, key P0 * prod I mod M, id 8-bit packed or
    colex rank; depth-first with indices ascending. Returns the exit status."""
    from . import lib as L
    env = os.environ if env is None else env
    try:
        q = parse_idump(text, env)
    except DumpError as e:
        err.write(e.message)
        return e.status
    n = len(q.INS)
    counts = np.zeros(n, np.uint64)
    total = 0
    if q.r <= n and q.Icap > 1:
        plan = L.Plan.insertion(q.INS, q.r, q.M, q.P0modM % q.M, q.Icap, min(q.split_idx, n), q.hmax,
                                L.CF_ID_COLEX if q.rank else L.CF_ID_PACK_ASC)
        try:
            counts = plan.first_counts()
            total = _write_plan(plan, range(n), out)
        finally:
            plan.close()
    out.flush()
    for t in range(q.T):
        err.write(f"t{t} {int(counts[t::q.T].sum()) if t < n else 0}\n")
    split_shown = q.split_idx
    err.write(f"TOTAL {total} records (r={q.r}, |INS|={n}, split_idx={split_shown}, hmax={q.hmax}, "
              f"ids={'colex-rank' if q.rank else '8-bit'})\n")
    return 0


def main(argv: Optional[List[str]] = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("ddump", "idump"):
        sys.stderr.write("usage: python -m cfjoin.dumpers ddump|idump  (input on stdin)\n")
        return 2
    text = sys.stdin.read()
    run = run_ddump if argv[0] == "ddump" else run_idump
    return run(text, sys.stdout.buffer, sys.stderr)


if __name__ == "__main__":
    sys.exit(main())
