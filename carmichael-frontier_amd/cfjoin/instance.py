"""Instance construction for the exchange-MITM residue join (host side, exact big-int).

An *instance* is one (modulus M, base B0, insertion pool INS, radius r) solve. Its D side is
every r-subset D of B0 with prod(D) >= Dmin; its I side is every r-subset I of INS with
prod(I) < Icap (exclusive) and at most `hmax` primes above SPLIT. A pair (D, I) joins when
    prod(D) == P0 * prod(I)   (mod M),      P0 = prod(B0).

The bounds follow gpu/mitm_gpu_mstar_driver.py:184-208 (run_instance) and :84-98 (hmax_for);
the pool follows gpu/run_mstar_family.py:35-40 (q <= cap) and gpu/mitm_gpu_mstar_driver.py:77-82
(q < cap); the family runner passes INS_PRIME_CAP = cap + 1, INS_MAX = 1024, SPLIT = 4000
(gpu/run_mstar_family.py:85-87).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from math import comb, prod
from typing import List, Optional, Tuple

from .verify import is_prime

K = 64
SPLIT_DEFAULT = 4000
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def pool(M: int, cap: int) -> List[int]:
    """Odd primes q <= cap with (q - 1) | M and q not dividing M (gpu/run_mstar_family.py:35-40)."""
    out = []
    q = 3
    while q <= cap:
        if M % (q - 1) == 0 and M % q != 0 and is_prime(q):
            out.append(q)
        q += 2
    return out


def hmax_for(B0: List[int], INS: List[int], r: int, U: int, split: int = SPLIT_DEFAULT) -> int:
    """Largest number h of insertion primes > split that can still give a completion below U;
    -1 if none (gpu/mitm_gpu_mstar_driver.py:84-98)."""
    small = [p for p in INS if p <= split]
    large = [p for p in INS if p > split]
    PB = prod(B0)
    Dmax = prod(sorted(B0)[-r:]) if r else 1
    best = -1
    for h in range(0, r + 1):
        if h > len(large) or r - h > len(small):
            continue
        v = PB * prod(small[: r - h]) * prod(large[:h])
        if v < U * Dmax:
            best = h
    return best


@dataclass
class Instance:
    """One exchange-MITM join instance, with every bound resolved to exact integers."""
    M: int
    r: int
    B0: List[int]                 # 64 base primes, ascending
    INS: List[int]                # insertion pool after the per-radius cap, ascending
    P0: int                       # prod(B0)
    Icap: int                     # exclusive bound on prod(I)
    Dmin: int                     # inclusive lower bound on prod(D)
    split: Optional[int]          # stratum split prime (None: no stratum)
    hmax: int                     # max insertion primes > split
    U: int = 0                    # the completion bound the instance was derived from
    n_ins_full: int = 0           # pool size before the per-radius cap
    tag: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def split_idx(self) -> int:
        """First index of INS whose prime exceeds split (len(INS) if none), as
        src/mitm64_idump_stream2.cpp:73 and gpu/igen_gpu.py:106."""
        if self.split is None:
            return len(self.INS)
        return next((i for i, p in enumerate(self.INS) if p > self.split), len(self.INS))

    @property
    def rank_ids(self) -> bool:
        """Colex-rank I ids are forced for r > 8 or pools > 256 (gpu/mitm_gpu_mstar_driver.py:206)."""
        return self.r > 8 or len(self.INS) > 256

    @property
    def host_exact(self) -> bool:
        """The reference routes tiny instances and Icap beyond 126 bits to a host join
        (gpu/mitm_gpu_mstar_driver.py:207)."""
        return self.Icap.bit_length() > 126 or comb(len(self.INS), self.r) < 200_000

    @property
    def P0modM(self) -> int:
        return self.P0 % self.M

    def descriptor(self) -> dict:
        return dict(M=self.M, r=self.r, nB=len(self.B0), nINS=len(self.INS), Icap=self.Icap,
                    Dmin=self.Dmin, split_idx=self.split_idx, hmax=self.hmax, tag=self.tag)


def build_instance(M: int, B0: List[int], INS_sorted: List[int], r: int, U: int,
                   split: Optional[int] = SPLIT_DEFAULT, hmax: Optional[int] = None,
                   tag: str = "") -> Optional[Instance]:
    """Resolve the bounds of one radius (gpu/mitm_gpu_mstar_driver.py:184-206).

    Returns None when the affordable pool holds fewer than r primes ("nothing to search",
    :204-205) or when no stratum can improve on U (hmax < 0, :309-310)."""
    B0s = sorted(B0)
    INS = list(INS_sorted)
    if hmax is None:
        hmax = hmax_for(B0s, INS, r, U, split if split is not None else 1 << 62)
    if hmax < 0:
        return None
    P0 = prod(B0s)
    topI = prod(INS[-r:])
    ptb = prod(B0s[-r:])
    Icap = max(1, min(topI + 1, (U * ptb) // P0 + 1))
    Imin = prod(INS[:r])
    Dmin = (P0 * Imin) // U + 1
    Imin1 = prod(INS[: r - 1])
    cap_r = (Icap - 1) // Imin1
    n_full = len(INS)
    INS = [q for q in INS if q <= cap_r]
    if len(INS) < r:
        return None
    return Instance(M=M, r=r, B0=B0s, INS=INS, P0=P0, Icap=Icap, Dmin=Dmin, split=split,
                    hmax=hmax, U=U, n_ins_full=n_full, tag=tag)


def family_instance(D: int, cap: int, U: int, r: int, ins_max: int = 1024,
                    split: int = SPLIT_DEFAULT, P: Optional[List[int]] = None) -> Optional[Instance]:
    """The ksmall instance of census modulus D at radius r, as the family runner sets it up
    (gpu/run_mstar_family.py:74-87 feeding gpu/mitm_gpu_mstar_driver.py:290-311)."""
    if P is None:
        P = pool(D, cap)
    B0 = P[:K]
    INS = sorted(p for p in P if p not in set(B0))[:ins_max]
    return build_instance(D, B0, INS, r, U, split=split, tag=f"D{D}_r{r}")


def load_workloads() -> dict:
    """Census rows and bounds of the two logged divisor families (data extracted from
    gpu/divisor_family_lambdaJ.json, gpu/mstar_divisor_family.json and the incumbent files)."""
    with open(os.path.join(_DATA, "workloads.json")) as f:
        return json.load(f)


def anchor_instance(which: str, r: int = 6) -> Instance:
    """The two flagship moduli at their family bound: 'M0' (1 768 248 177 696 000, U = N0) and
    'lambdaN' (155 016 423 578 016 000, U = N)."""
    w = load_workloads()
    fam = w["families"][which]
    row = next(x for x in fam["rows"] if int(x["D"]) == int(fam["anchor"]))
    inst = family_instance(int(row["D"]), int(row["cap"]), int(fam["U"]), r)
    assert inst is not None
    inst.tag = f"{which}_r{r}"
    return inst


def c1_modulus(seed: int) -> int:
    """Plain k=29 workload modulus: M = 2^a 3^b 5^c 7 11 13 with a in [4,7], b in [2,4], c in [1,2]
    drawn from random.Random(seed)."""
    import random
    g = random.Random(seed)
    a = g.randint(4, 7)
    b = g.randint(2, 4)
    c = g.randint(1, 2)
    return 2 ** a * 3 ** b * 5 ** c * 7 * 11 * 13


def c1_instance(seed: int, s: Optional[int] = None, r: int = 3, k: int = 29, qmax: int = 20000) -> Instance:
    """The k = 29, r = 3 single-modulus workload. Q(M) = odd primes q <= qmax with (q-1) | M, q not
    dividing M; B0 = the k smallest; the bound is U = s * prod(B0) through the same recipe as the
    k = 64 instances (s = None: I side unbounded, every r-subset of the pool; D side unfiltered)."""
    M = c1_modulus(seed)
    Q = pool(M, qmax)
    B0 = Q[:k]
    INS = Q[k:]
    P0 = prod(B0)
    if s is None:
        return Instance(M=M, r=r, B0=B0, INS=INS, P0=P0, Icap=prod(INS[-r:]) + 1, Dmin=1, split=None,
                        hmax=r, U=0, n_ins_full=len(INS), tag=f"C1_seed{seed}_unbounded")
    inst = build_instance(M, B0, INS, r, s * P0, split=None, hmax=r, tag=f"C1_seed{seed}_s{s}")
    if inst is None:  # empty instance: keep it as the zero-record edge case
        return Instance(M=M, r=r, B0=B0, INS=INS, P0=P0, Icap=1, Dmin=prod(B0[-r:]) + 1, split=None,
                        hmax=r, U=s * P0, n_ins_full=len(INS), tag=f"C1_seed{seed}_s{s}_empty")
    return inst


def c2_modulus(seed: int, pool_lo: int = 200, pool_hi: int = 500) -> Tuple[int, int]:
    """Plain k=64 workload modulus: M = 2^a 3^b 5^c 7^d times a seeded subset of the primes 11..61,
    drawn with random.Random(seed) until M lies in [1.7e15, 1.6e17]; then the pool cap is raised from
    2000 until Q(M) (primes q <= cap with (q-1) | M, q not dividing M) holds pool_lo..pool_hi primes.
    Returns (M, cap)."""
    import random
    g = random.Random(seed)
    small = [11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61]
    for _ in range(100000):
        M = 2 ** g.randint(5, 9) * 3 ** g.randint(2, 4) * 5 ** g.randint(1, 3) * 7 ** g.randint(1, 2)
        for q in small:
            if g.random() < 0.5:
                M *= q
        if not 1.7e15 <= M <= 1.6e17:
            continue
        lo, hi = 2000, 2000
        while len(pool(M, hi)) < pool_lo and hi < (1 << 24):
            lo, hi = hi, hi * 2
        if len(pool(M, hi)) < pool_lo:
            continue
        while hi - lo > 1:                   # smallest cap reaching pool_lo primes
            mid = (lo + hi) // 2
            if len(pool(M, mid)) >= pool_lo:
                hi = mid
            else:
                lo = mid
        if len(pool(M, hi)) <= pool_hi:
            return M, hi
    raise ValueError(f"no k=64 modulus found for seed {seed}")


def c2_instance(seed: int, s: int, r: int = 6, k: int = 64) -> Optional[Instance]:
    """The k = 64, r = 6 single-modulus workload on a seeded modulus: Q(M) as in c2_modulus (200-500
    primes), B0 = the 64 smallest, bound U = s * prod(B0) through the same recipe as the family
    instances (the literal 1.1 * prod(B0) leaves both sides empty)."""
    M, cap = c2_modulus(seed)
    Q = pool(M, cap)
    B0 = Q[:k]
    INS = Q[k:]
    return build_instance(M, B0, INS, r, s * prod(B0), split=None, hmax=r, tag=f"C2_seed{seed}_s{s}")


# ---- positive controls: instances that must contain a known k = 64 Carmichael number ---------------
CONTROLS = {
    "N": "results_k64_best_global.json",                           # lambda(N) = 155016423578016000, r = 6
    "N_prev": "results_k64_best_global_prev_148_N.json",           # M0 = 1768248177696000, r = 7
    "N0": "results_k64_best_global_prev_148_539839443357.json",    # M0, r = 8
    "N149": "results_k64.json",                                    # 353649635539200, r = 9
}


@dataclass
class Control:
    """A positive control: the instance, the incumbent n it must reproduce and the pair that gives it."""
    name: str
    N: int
    factors: List[int]
    D_need: List[int]             # base primes n leaves out
    I_need: List[int]             # pool primes n brings in
    inst: Instance


def control_instance(name: str, nins: Optional[int] = None, pool_below: int = 4000) -> Control:
    """The incumbent's own modulus with the ksmall base (the 64 smallest pool primes below
    pool_below) and an insertion pool holding the incumbent's inserted primes, so that the join
    must emit the pair (D_need, I_need) whose completion is the incumbent
    (gpu/validate_mitm_gpu2_pos.py:20-36). nins: the pool is I_need plus the first nins - r other
    pool primes, as the validator builds it; None: every pool prime outside the base. The bound is
    U = n + 1, so the instance's product bounds admit n itself; no stratum split."""
    inc = load_workloads()["incumbents"][CONTROLS[name]]
    M, N = int(inc["modulus"]), int(inc["n"])
    fac = sorted(int(p) for p in inc["factors"])
    P = pool(M, pool_below - 1)
    B0 = P[:K]
    in_base = set(B0)
    D_need = sorted(in_base - set(fac))
    I_need = sorted(set(fac) - in_base)
    r = len(D_need)
    assert r == len(I_need) and set(I_need) <= set(P), "the incumbent is not reachable from the ksmall base"
    rest = [p for p in P if p not in in_base and p not in set(I_need)]
    if nins is not None:
        rest = rest[: max(nins - r, 0)]
    inst = build_instance(M, B0, sorted(I_need + rest), r, N + 1, split=None, hmax=r,
                          tag=f"control_{name}_r{r}" + ("" if nins is None else f"_ins{nins}"))
    assert inst is not None and set(I_need) <= set(inst.INS)
    return Control(name, N, fac, D_need, I_need, inst)
