#!/usr/bin/env python3
"""igen_gpu.py -- drop-in for the reference module of the same name (gpu/igen_gpu.py): the
filtered insertion side of the exchange MITM, generated on MI355X by libcfjoin.

Emits exactly the records of src/mitm64_idump_stream2 with IDUMP_RANK=1: one (key, id) per
r-subset I of the ascending pool INS with prod(I) < Icap and at most HMAX primes > SPLIT,
    key = (P0 * prod(I)) mod M,     id = colex rank sum_j C(i_j, j+1).
Same functions and arguments as the reference:
  quotient_table(INS, r, Icap)                                      gpu/igen_gpu.py:91-100
  generate_chunks(INS, r, M, P0, Icap, split, hmax, max_records, block_bytes)   :113-145
  generate(INS, r, M, P0, Icap, split, hmax) -> (keys, ids, count)  :147-152
Chunks cover contiguous first-index ranges in first-index order and are cut exactly where the
reference cuts them (first-index blocks of block_bytes // (8 n^2), then greedy by max_records);
within a chunk, records are in depth-first (lexicographic) order. The generator's return value
(StopIteration.value) is the total record count.

Here the whole enumeration tree is expanded level by level on the device (count -> scan -> write
per level), so there is no per-thread DFS and no load imbalance at r = 7, 8; the count per
first index comes from the same expansion.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from cfjoin import lib as _lib  # noqa: E402
from cfjoin import xp as cp  # noqa: E402


def quotient_table(INS, r, Icap):
    """H[t][x] = floor((Icap-1) / (a_x ... a_{x+t})) for 0 <= t <= r-1 and x + t < n (0 otherwise); exact.
    Window products come from exact prefix products: a_x ... a_{x+t} = pre[x+t+1] / pre[x]."""
    n = len(INS)
    pre = [1]
    for a in INS:
        pre.append(pre[-1] * int(a))
    top = Icap - 1
    return [top // (pre[x + t + 1] // pre[x]) if x + t < n else 0 for t in range(r) for x in range(n)]

def _prepare(INS, r, M, P0, Icap, split, hmax):
    n = len(INS)
    assert 3 <= r <= 8 and n <= 1024 and 1 <= M < (1 << 64)
    assert all(INS[i] < INS[i + 1] for i in range(n - 1)), "INS must be strictly ascending"
    assert Icap - 1 < (1 << 128), "Icap needs more than 128 bits: use the CPU dumper"
    split_idx = n if split is None else next((i for i, p in enumerate(INS) if p > split), n)
    hmax = r if hmax is None else hmax
    assert comb_fits(n, r), "rank ids would overflow u64"
    plan = _lib.Plan.insertion(INS, r, M, P0 % M, Icap, split_idx, hmax, _lib.CF_ID_COLEX)
    return n, plan

def comb_fits(n, r):
    from math import comb
    return comb(n, r) < 2 ** 64

def generate_chunks(INS, r, M, P0, Icap, split=None, hmax=None, max_records=200_000_000, block_bytes=512 << 20):
    """Yield (keys, ids) device uint64 chunks covering exactly the filtered r-subsets of INS; the
    generator's return value (StopIteration.value) is the total record count. First indices are
    processed in blocks [a, b) with (b-a)*n*n*8 <= block_bytes; within a block, first indices are
    grouped greedily while the chunk stays <= max_records (a single first index may exceed it)."""
    if Icap <= 1:
        return 0
    n, plan = _prepare(INS, r, M, P0, Icap, split, hmax)
    try:
        per_first = [int(c) for c in plan.first_counts()]
        ib = max(1, min(n, block_bytes // (8 * n * n)))
        total = 0
        a = 0
        while a < n:
            b = min(n, a + ib)
            a2 = a
            while a2 < b:
                b2 = a2 + 1
                acc = per_first[a2]
                while b2 < b and acc + per_first[b2] <= max_records:
                    acc += per_first[b2]
                    b2 += 1
                if acc:
                    keys = cp.empty(acc, dtype=cp.uint64)
                    ids = cp.empty(acc, dtype=cp.uint64)
                    wrote = plan.write(a2, b2, keys.ptr, ids.ptr)
                    assert wrote == acc, ("per-prefix count/write mismatch", a2, b2)
                    keys.key_bound = M
                    total += acc
                    yield keys, ids
                a2 = b2
            a = b
        return total
    finally:
        plan.close()

def generate(INS, r, M, P0, Icap, split=None, hmax=None):
    """Single-shot convenience: concatenated (keys, ids, count)."""
    ks, is_ = [], []
    for k, i in generate_chunks(INS, r, M, P0, Icap, split, hmax, max_records=1 << 62):
        ks.append(k)
        is_.append(i)
    if not ks:
        return cp.empty(0, cp.uint64), cp.empty(0, cp.uint64), 0
    if len(ks) == 1:
        return ks[0], is_[0], int(ks[0].size)
    k = cp.concatenate(ks)
    i = cp.concatenate(is_)
    return k, i, int(k.size)
