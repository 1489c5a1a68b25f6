#!/usr/bin/env python3
"""mitm_gpu2.py -- drop-in for the reference module of the same name (gpu/mitm_gpu2.py), running
on MI355X through libcfjoin's hand-written gfx950 kernels.

Same functions, same arguments, same return meaning and the same AssertionError preconditions:
  generate(primes, r, M, mult=1, bits=10, lo=0, hi=None) -> (keys, ids)   gpu/mitm_gpu2.py:79-98
  sort_side(keys, ids) -> (keys, ids)                                    gpu/mitm_gpu2.py:101-104
  join_sorted(Dk, Did, Ik, Iid) -> (Did_pairs, Iid_pairs)                gpu/mitm_gpu2.py:107-119
  decode_ids(ids, r, bits) -> [tuple]                                    gpu/mitm_gpu2.py:122-126
Arrays are cfjoin.xp.ndarray (device, uint64). One deliberate widening: any M < 2^64 is accepted
(the reference asserted M < 2^62 here but M < 2^64 in igen_gpu.py:103); keys stay canonical
residues in [0, M), so equal key still means equal residue.
"""
import os
import sys
from math import comb

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from cfjoin import lib as _lib  # noqa: E402
from cfjoin import xp as cp  # noqa: E402


def generate(primes, r: int, M: int, mult: int = 1, bits: int = 10, lo: int = 0, hi=None):
    """Residue records for the r-combinations of `primes` with CNS rank in [lo, hi).
    Returns device uint64 (keys, ids); ids pack r indices of `bits` bits each."""
    n = len(primes)
    assert 1 <= r <= 10 and r * bits <= 64, f"r={r}, bits={bits}: need r*bits<=64, r<=10"
    assert n <= (1 << bits), f"pool {n} does not fit {bits}-bit indices"
    assert 1 <= M < (1 << 64), "M must be a uint64 so the residue key fits u64 uniquely"
    total = comb(n, r)
    if hi is None:
        hi = total
    assert 0 <= lo <= hi <= total, (lo, hi, total)
    count = hi - lo
    keys = cp.empty(count, dtype=cp.uint64)
    ids = cp.empty(count, dtype=cp.uint64)
    if count == 0:
        return keys, ids
    pr = _lib.u64_array(primes)
    ctx = keys.ctx
    _lib.check(_lib.load().cf_gen_subsets(ctx.handle, pr.ctypes.data_as(_lib.c_void_p), n, r, bits, M, mult % (1 << 64),
                                          lo, count, _lib.c_void_p(keys.ptr), _lib.c_void_p(ids.ptr)), "generate")
    keys.key_bound = M
    return keys, ids


def sort_side(keys, ids):
    """Sort one side by key (stable). Returns new (sorted_keys, sorted_ids); inputs are unchanged."""
    if keys.size <= 1:
        return keys.copy(), ids.copy()
    k = cp.ndarray(keys.size, keys.dtype, ctx=keys.ctx)
    i = cp.ndarray(ids.size, ids.dtype, ctx=keys.ctx)
    # keys made by this library's generators carry their modulus: no pass for the largest key
    bound = keys.key_bound or 0
    _lib.check(_lib.load().cf_sort_pairs_into(keys.ctx.handle, _lib.c_void_p(keys.ptr), _lib.c_void_p(ids.ptr),
                                              _lib.c_void_p(k.ptr), _lib.c_void_p(i.ptr), keys.size, bound), "sort_side")
    k.key_bound = keys.key_bound
    return k, i


def join_sorted(Dk, Did, Ik, Iid):
    """Sort-merge join of two ALREADY-SORTED sides. Emits EVERY equal-key (D,I) pair."""
    ctx = Dk.ctx
    od, oi = _lib.c_void_p(), _lib.c_void_p()
    n = _lib.c_u64()
    _lib.check(_lib.load().cf_join_sorted(ctx.handle, _lib.c_void_p(Dk.ptr), _lib.c_void_p(Did.ptr), Dk.size,
                                          _lib.c_void_p(Ik.ptr), _lib.c_void_p(Iid.ptr), Ik.size,
                                          _lib.ctypes.byref(od), _lib.ctypes.byref(oi), _lib.ctypes.byref(n)),
               "join_sorted")
    if n.value == 0:
        return cp.empty(0, dtype=cp.uint64), cp.empty(0, dtype=cp.uint64)
    return cp.ndarray._adopt(od.value, n.value, ctx=ctx), cp.ndarray._adopt(oi.value, n.value, ctx=ctx)


def decode_ids(ids, r: int, bits: int):
    """Unpack ids -> list of index tuples (host): field p of an id is bits [p*bits, (p+1)*bits)."""
    host = np.asarray(ids.get() if isinstance(ids, cp.ndarray) else ids, dtype=np.uint64).reshape(-1)
    shifts = np.arange(r, dtype=np.uint64) * np.uint64(bits)
    fields = (host[:, None] >> shifts[None, :]) & np.uint64((1 << bits) - 1)
    return [tuple(int(x) for x in row) for row in fields]
