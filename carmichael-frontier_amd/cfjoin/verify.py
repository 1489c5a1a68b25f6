"""Host re-check of join hits: exact completion and Korselt's criterion (exact big-int).

The certificate check answers the question the reference's ref/ref_carmichael.py:55-72 answers,
with the same result tuple: is n = prod(primes) a Carmichael number with exactly len(primes)
distinct prime factors? It is written from the textbook statements:
  * Korselt (1899): a composite n is Carmichael iff n is squarefree and (p - 1) | (n - 1) for
    every prime p | n; with at least three distinct prime factors n is composite and squarefree
    by construction, so only the divisibility conditions remain;
  * the strong-probable-prime (Miller-Rabin) test: write n - 1 = d * 2^s with d odd; n passes base
    a if a^d = 1 or a^(d 2^j) = -1 (mod n) for some 0 <= j < s. Deterministic base sets are known:
    Sinclair's seven bases for n < 2^64, and the first twelve primes for n < 3.3e24 (Sorenson and
    Webster, 2015); larger n get both sets, a probabilistic answer as in the reference.
The per-pair completion follows gpu/mitm_gpu_mstar_driver.py:226-237 (S = (B0 minus D) union I,
n = P0 * prod(I) / prod(D)).
"""
from __future__ import annotations

from math import comb, prod
from typing import Iterable, List, Sequence, Tuple

_SINCLAIR_BASES = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)   # every n < 2^64
_PRIME_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)             # every n < 3.317e24
_SMALL = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47)


def _strong_probable_prime(n: int, a: int, d: int, s: int) -> bool:
    """n (odd, > 2) passes the strong test to base a, with n - 1 = d * 2^s, d odd."""
    a %= n
    if a == 0:          # a base that is a multiple of n says nothing
        return True
    x = pow(a, d, n)
    if x == 1 or x == n - 1:
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def is_prime(n: int) -> bool:
    """Primality of n by trial division by the primes below 50, then strong tests to a base set
    that is deterministic below 3.3e24."""
    if n < 2:
        return False
    for p in _SMALL:
        if n % p == 0:
            return n == p
    if n < 53 * 53:
        return True
    s = ((n - 1) & -(n - 1)).bit_length() - 1      # the power of two in n - 1
    d = (n - 1) >> s
    bases = _SINCLAIR_BASES if n < (1 << 64) else _PRIME_BASES + _SINCLAIR_BASES
    return all(_strong_probable_prime(n, a, d, s) for a in bases)


def verify_certificate(primes: Sequence[int]) -> Tuple[bool, int, str]:
    """(ok, n, reason) for the claim that n = prod(primes) is Carmichael with exactly
    len(primes) distinct prime factors. n is reported as 0 when the factor list itself is
    invalid (too short, repeated or non-prime entries), as the reference does."""
    ps = [int(p) for p in primes]
    k = len(ps)
    if k < 3:
        return (False, 0, f"only {k} factors: Korselt needs at least three distinct primes")
    if len(set(ps)) < k:
        return (False, 0, "a factor is repeated, so n is not squarefree")
    composite = [p for p in ps if not is_prime(p)]
    if composite:
        return (False, 0, f"factor {composite[0]} is not prime")
    n = prod(ps)
    for p in ps:
        if (n - 1) % (p - 1):
            return (False, n, f"p = {p}: p - 1 does not divide n - 1")
    return (True, n, f"Carmichael, {k} distinct prime factors")


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
