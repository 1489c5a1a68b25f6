"""ctypes wrapper of liboracle.so, the CPU restatement of the residue join (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg of
bench.py. The product (carmichael-frontier_amd/) never imports it.

Every function takes exact Python ints and numpy uint64 arrays and returns numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
_lib = None

u64 = ctypes.c_uint64
P = ctypes.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.or_gen_rank.restype = ctypes.c_longlong
        L.or_gen_rank.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, u64, u64, u64, u64, P, P]
        L.or_ddump.restype = ctypes.c_longlong
        L.or_ddump.argtypes = [u64, ctypes.c_int, u64, u64, P, ctypes.c_int, ctypes.c_int, P, P, u64]
        L.or_idump.restype = ctypes.c_longlong
        L.or_idump.argtypes = [u64, ctypes.c_int, u64, u64, u64, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, u64]
        L.or_join_sorted.restype = ctypes.c_longlong
        L.or_join_sorted.argtypes = [P, P, u64, P, P, u64, P, P, u64]
        L.or_cpu_join.restype = ctypes.c_longlong
        L.or_cpu_join.argtypes = [u64, ctypes.c_int, u64, u64, u64, u64, u64, P, ctypes.c_int, P, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, u64, P, P]
        L.or_join_dsweep.restype = ctypes.c_longlong
        L.or_join_dsweep.argtypes = L.or_cpu_join.argtypes
        _lib = L
    return _lib


def _arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _split128(v: int) -> Tuple[int, int]:
    if v < 0 or v >= 1 << 128:
        raise ValueError("value must fit 128 bits")
    return v & ((1 << 64) - 1), v >> 64


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise AssertionError(f"{what}: oracle returned error {rc}")
    return rc


def gen_rank(primes, r: int, M: int, mult: int = 1, bits: int = 10, lo: int = 0, count: Optional[int] = None):
    from math import comb
    pr = _arr(primes)
    if count is None:
        count = comb(len(pr), r) - lo
    k = np.empty(count, np.uint64)
    i = np.empty(count, np.uint64)
    _check(lib().or_gen_rank(_ptr(pr), len(pr), r, bits, M, mult, lo, count, _ptr(k), _ptr(i)), "gen_rank")
    return k, i


def ddump(M: int, r: int, Dmin: int, B, bitmask: bool = False):
    b = _arr(B)
    lo, hi = _split128(Dmin)
    n = _check(lib().or_ddump(M, r, lo, hi, _ptr(b), len(b), int(bitmask), None, None, 0), "ddump")
    k = np.empty(n, np.uint64)
    i = np.empty(n, np.uint64)
    lib().or_ddump(M, r, lo, hi, _ptr(b), len(b), int(bitmask), _ptr(k), _ptr(i), n)
    return k, i


def idump(M: int, r: int, P0modM: int, Icap: int, INS, split_idx: Optional[int] = None,
          hmax: Optional[int] = None, rank: bool = True, i_lo: int = 0, i_hi: Optional[int] = None):
    ins = _arr(INS)
    n = len(ins)
    if split_idx is None:
        split_idx = 1 << 30
    if hmax is None:
        hmax = 1 << 30
    if i_hi is None:
        i_hi = n
    lo, hi = _split128(Icap)
    args = (M, r, P0modM, lo, hi, _ptr(ins), n, split_idx, hmax, int(rank), i_lo, i_hi)
    cnt = _check(lib().or_idump(*args, None, None, 0), "idump")
    k = np.empty(cnt, np.uint64)
    i = np.empty(cnt, np.uint64)
    lib().or_idump(*args, _ptr(k), _ptr(i), cnt)
    return k, i


def join_sorted(Dk, Did, Ik, Iid):
    Dk, Did, Ik, Iid = map(_arr, (Dk, Did, Ik, Iid))
    n = _check(lib().or_join_sorted(_ptr(Dk), _ptr(Did), len(Dk), _ptr(Ik), _ptr(Iid), len(Ik), None, None, 0), "join")
    od = np.empty(n, np.uint64)
    oi = np.empty(n, np.uint64)
    lib().or_join_sorted(_ptr(Dk), _ptr(Did), len(Dk), _ptr(Ik), _ptr(Iid), len(Ik), _ptr(od), _ptr(oi), n)
    return od, oi


def cpu_join(M: int, r: int, P0modM: int, Icap: int, Dmin: int, B, INS, split_idx: Optional[int] = None,
             hmax: Optional[int] = None, rank: bool = True, threads: int = 1, cap: int = 1 << 20):
    """Hash join in the reference CPU engine's shape. Returns dict(pairs_d, pairs_i, nD, nI, td, ti)."""
    b = _arr(B)
    ins = _arr(INS)
    if split_idx is None:
        split_idx = 1 << 30
    if hmax is None:
        hmax = 1 << 30
    il, ih = _split128(Icap)
    dl, dh = _split128(Dmin)
    od = np.empty(cap, np.uint64)
    oi = np.empty(cap, np.uint64)
    stats = np.zeros(3, np.uint64)
    times = np.zeros(2, np.float64)
    n = _check(lib().or_cpu_join(M, r, P0modM, il, ih, dl, dh, _ptr(b), len(b), _ptr(ins), len(ins), split_idx,
                                 hmax, int(rank), threads, _ptr(od), _ptr(oi), cap, _ptr(stats),
                                 times.ctypes.data_as(ctypes.c_void_p)), "cpu_join")
    if n > cap:
        raise AssertionError("pair buffer too small")
    return dict(pairs_d=od[:n].copy(), pairs_i=oi[:n].copy(), nD=int(stats[0]), nI=int(stats[1]),
                td=float(times[0]), ti=float(times[1]))


def join_dsweep(M: int, r: int, P0modM: int, Icap: int, Dmin: int, B, INS, split_idx: Optional[int] = None,
                hmax: Optional[int] = None, rank: bool = True, threads: int = 1, cap: int = 1 << 20):
    """The CPU join with the roles swapped (hash the I side, sweep the D side on `threads` threads),
    for D sides far larger than the I side. Same pair set as cpu_join. Returns dict(pairs_d, pairs_i,
    nD, nI, ti (I hash s), td (D sweep s))."""
    b = _arr(B)
    ins = _arr(INS)
    split_idx = 1 << 30 if split_idx is None else split_idx
    hmax = 1 << 30 if hmax is None else hmax
    il, ih = _split128(Icap)
    dl, dh = _split128(Dmin)
    od = np.empty(cap, np.uint64)
    oi = np.empty(cap, np.uint64)
    stats = np.zeros(3, np.uint64)
    times = np.zeros(2, np.float64)
    n = _check(lib().or_join_dsweep(M, r, P0modM, il, ih, dl, dh, _ptr(b), len(b), _ptr(ins), len(ins), split_idx,
                                    hmax, int(rank), threads, _ptr(od), _ptr(oi), cap, _ptr(stats),
                                    times.ctypes.data_as(ctypes.c_void_p)), "join_dsweep")
    if n > cap:
        raise AssertionError("pair buffer too small")
    return dict(pairs_d=od[:n].copy(), pairs_i=oi[:n].copy(), nD=int(stats[0]), nI=int(stats[1]),
                ti=float(times[0]), td=float(times[1]))


def exact_join(M: int, r: int, P0: int, Icap: int, Dmin: int, B, INS, split_idx: Optional[int] = None,
               hmax: Optional[int] = None):
    """The join in Python integers, for moduli of any size (small instances only): every pair of an
    r-subset D of B with prod(D) >= Dmin and an r-subset I of INS with prod(I) < Icap and at most hmax
    indices >= split_idx, with prod(D) = P0 * prod(I) (mod M). Follows the record sets of
    src/mitm64_ddump.cpp:29-44 (D chosen from the largest base prime down, cut once even the largest
    completion stays below Dmin) and src/mitm64_idump_stream2.cpp:47-57 (I chosen upward, cut once the
    cheapest completion reaches Icap), joined by residue as src/mitm64.cpp:51-121 does.
    Returns (nD, nI, set of (D index tuple, I index tuple)), indices ascending into B and INS."""
    B = [int(x) for x in B]
    INS = [int(x) for x in INS]
    split_idx = len(INS) if split_idx is None else split_idx
    hmax = r if hmax is None else hmax
    by_res = {}
    nD = 0

    def drec(hi, t, prod, chosen):
        nonlocal nD
        if t == 0:
            nD += 1
            by_res.setdefault(prod % M, []).append(tuple(sorted(chosen)))
            return
        for i in range(hi, t - 2, -1):
            best = prod
            for q in range(i, i - t, -1):
                best *= B[q]
            if best < Dmin:      # even the largest completion from here is too small, and it only shrinks
                break
            drec(i - 1, t - 1, prod * B[i], chosen + [i])

    drec(len(B) - 1, r, 1, [])
    pairs = set()
    nI = 0
    P0m = P0 % M

    def irec(lo, t, prod, chosen, nlarge):
        nonlocal nI
        if t == 0:
            nI += 1
            for d in by_res.get(P0m * prod % M, ()):
                pairs.add((d, tuple(chosen)))
            return
        for i in range(lo, len(INS) - t + 1):
            cheap = prod
            for q in range(i, i + t):
                cheap *= INS[q]
            if cheap >= Icap:    # the cheapest completion from here is too large, and it only grows
                break
            nl = nlarge + (i >= split_idx)
            if nl > hmax:
                continue
            irec(i + 1, t - 1, prod * INS[i], chosen + [i], nl)

    irec(0, r, 1, [], 0)
    return nD, nI, pairs


def instance_records(inst, rank: Optional[bool] = None):
    """(Dk, Did, Ik, Iid) of an Instance (cfjoin.instance.Instance-like object)."""
    Dk, Did = ddump(inst.M, inst.r, inst.Dmin, inst.B0, bitmask=inst.r > 10)
    Ik, Iid = idump(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.INS, inst.split_idx, inst.hmax,
                    rank=inst.rank_ids if rank is None else rank)
    return Dk, Did, Ik, Iid


def pair_set(pd, pi) -> np.ndarray:
    """Canonical form of a pair multiset: (D id, I id) rows sorted lexicographically."""
    a = np.stack([np.asarray(pd, np.uint64), np.asarray(pi, np.uint64)], axis=1) if len(pd) else np.zeros((0, 2), np.uint64)
    if len(a):
        a = a[np.lexsort((a[:, 1], a[:, 0]))]
    return a


# ---- the compiled reference tools (oracle/_ref), present only where they were built ---------

def ref_available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, b)) for b in ("mitm64", "mitm64_ddump", "mitm64_idump_stream2"))


def ref_ddump(M: int, r: int, Dmin: int, B, bitmask: bool = False) -> np.ndarray:
    env = dict(os.environ, DDUMP_BITMASK="1" if bitmask else "0")
    inp = f"{M} {r} {Dmin}\n{len(B)} " + " ".join(map(str, B)) + "\n"
    pr = subprocess.run([os.path.join(REF_DIR, "mitm64_ddump")], input=inp.encode(), capture_output=True, check=True, env=env)
    return np.frombuffer(pr.stdout, dtype=np.uint64).reshape(-1, 2)


def ref_idump(M: int, r: int, P0modM: int, Icap: int, INS, split: Optional[int], hmax: Optional[int],
              rank: bool, threads: int = 1) -> np.ndarray:
    env = dict(os.environ, IDUMP_RANK="1" if rank else "0")
    if split is not None:
        env["IDUMP_SPLIT"] = str(split)
    if hmax is not None:
        env["IDUMP_HMAX"] = str(hmax)
    inp = "\n".join([f"{M} {r} {P0modM} {Icap} 1 {threads} 0", "0", f"{len(INS)} " + " ".join(map(str, INS))]) + "\n"
    pr = subprocess.run([os.path.join(REF_DIR, "mitm64_idump_stream2")], input=inp.encode(), capture_output=True,
                        check=True, env=env)
    return np.frombuffer(pr.stdout, dtype=np.uint64).reshape(-1, 2)


def ref_mitm64(M: int, r: int, P0modM: int, Icap_incl: int, Dmin: int, B, INS, threads: int, maxmatch: int = 1 << 30):
    """Runs the reference CPU join; returns (matches [(D primes, I primes)], stats dict from stderr)."""
    inp = "\n".join([f"{M} {r} {P0modM} {Icap_incl} {Dmin} {threads} {maxmatch}", f"{len(B)} " + " ".join(map(str, B)),
                     f"{len(INS)} " + " ".join(map(str, INS))]) + "\n"
    pr = subprocess.run([os.path.join(REF_DIR, "mitm64")], input=inp.encode(), capture_output=True, check=True)
    out = []
    for line in pr.stdout.decode().splitlines():
        if line.startswith("MATCH"):
            parts = line.split("|")
            out.append((tuple(sorted(int(x) for x in parts[1].split()[1:])), tuple(sorted(int(x) for x in parts[2].split()[1:]))))
    err = pr.stderr.decode()
    import re
    m = re.search(r"D_leaves=(\d+) Dres=(\d+) \(([\d.]+)s\)\s+I_leaves=(\d+) matches=(\d+) kept=(\d+) \(([\d.]+)s\)", err)
    st = dict(D_leaves=int(m.group(1)), Dres=int(m.group(2)), td=float(m.group(3)), I_leaves=int(m.group(4)),
              matches=int(m.group(5)), ti=float(m.group(7))) if m else {"stderr": err}
    return out, st
