"""Host re-check of join hits: exact completion and Korselt's criterion (exact big-int).

Restates the certificate check of ref/ref_carmichael.py:30-72 (deterministic Miller-Rabin with
the first twelve primes as witnesses, then Korselt: n squarefree with >= 3 prime factors and
(p - 1) | (n - 1) for every p | n) and the per-pair completion of
gpu/mitm_gpu_mstar_driver.py:226-237 (S = (B0 minus D) union I, n = P0 * prod(I) / prod(D)).
"""
from __future__ import annotations

from math import comb, prod
from typing import Iterable, List, Sequence, Tuple

_WITNESSES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Strong-probable-prime test to the twelve smallest prime bases; deterministic below
    3.3e24 (ref/ref_carmichael.py:30-52)."""
    if n < 2:
        return False
    for p in _WITNESSES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _WITNESSES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def verify_certificate(primes: Sequence[int]) -> Tuple[bool, int, str]:
    """(ok, n, reason) for the claim that n = prod(primes) is Carmichael with exactly
    len(primes) prime factors (ref/ref_carmichael.py:55-72)."""
    primes = list(primes)
    if len(primes) < 3:
        return (False, 0, "a Carmichael number has at least 3 prime factors")
    if len(set(primes)) != len(primes):
        return (False, 0, "factors are not distinct (n must be squarefree)")
    for p in primes:
        if not is_prime(p):
            return (False, 0, f"claimed factor {p} is not prime")
    n = prod(primes)
    for p in primes:
        if (n - 1) % (p - 1):
            return (False, n, f"Korselt fails: (p-1)={p - 1} does not divide n-1 for p={p}")
    return (True, n, "Carmichael with exactly %d prime factors" % len(primes))


def decode_packed(v: int, r: int, bits: int) -> Tuple[int, ...]:
    """Unpack r fields of `bits` bits, lowest field first (gpu/mitm_gpu2.py:122-126)."""
    mask = (1 << bits) - 1
    return tuple((int(v) >> (p * bits)) & mask for p in range(r))


def decode_bitmask(v: int) -> Tuple[int, ...]:
    """Indices of the set bits of a 64-bit deletion mask (src/mitm64_ddump.cpp:32)."""
    v = int(v)
    return tuple(b for b in range(64) if (v >> b) & 1)


def unrank_colex(v: int, r: int, n: int) -> Tuple[int, ...]:
    """Inverse of id = sum_j C(i_j, j + 1) over i_0 < ... < i_{r-1} < n
    (gpu/mitm_gpu_mstar_driver.py:120-126)."""
    v = int(v)
    idx = []
    x = n - 1
    for j in range(r - 1, -1, -1):
        while comb(x, j + 1) > v:
            x -= 1
        idx.append(x)
        v -= comb(x, j + 1)
        x -= 1
    if v != 0:
        raise ValueError("not a colex rank of an r-subset")
    return tuple(reversed(idx))


def completion(B0: Sequence[int], INS: Sequence[int], d_idx: Iterable[int], i_idx: Iterable[int],
               P0: int, M: int):
    """Exact completion of one pair: returns (factors, n) or None when the pair is not an
    integral completion. Also asserts the residue identity the join claims."""
    D = [B0[x] for x in d_idx]
    I = [INS[x] for x in i_idx]
    prodD, prodI = prod(D), prod(I)
    if (P0 * prodI - prodD) % M:
        raise AssertionError("joined pair has unequal residues mod M")
    if (P0 * prodI) % prodD:
        return None
    n = P0 * prodI // prodD
    fac = sorted((set(B0) - set(D)) | set(I))
    return fac, n


def recheck_pairs(inst, d_ids, i_ids, d_decode, i_decode) -> List[dict]:
    """Decode every pair, rebuild the 64-prime completion exactly and run the certificate
    check on it (gpu/mitm_gpu_mstar_driver.py:260-273). Returns one dict per pair."""
    out = []
    for dv, iv in zip(d_ids, i_ids):
        d = d_decode(int(dv))
        i = i_decode(int(iv))
        c = completion(inst.B0, inst.INS, d, i, inst.P0, inst.M)
        rec = {"D": [inst.B0[x] for x in d], "I": [inst.INS[x] for x in i], "integral": c is not None}
        if c is not None:
            fac, n = c
            ok, _, why = verify_certificate(fac) if len(fac) == len(inst.B0) else (False, n, "not 64 distinct")
            rec.update(n_digits=len(str(n)), carmichael=ok, below_U=bool(ok and inst.U and n < inst.U), reason=why)
        out.append(rec)
    return out
