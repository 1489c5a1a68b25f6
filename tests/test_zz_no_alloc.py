"""cf_join_sorted_noalloc against cf_join_sorted, through the C ABI.

Two calls in a row, the second with a smaller input (the scratch keeps the first call's contents,
so a stale count or offset would show), each compared with the allocating entry point pair for pair;
the context pool's live and cached bytes are unchanged by the no-allocation calls; a capacity below
the pair count and an I side above the limit are rejected with CF_EINVAL.
"""
import ctypes

import numpy as np
import pytest

from cfjoin import lib as L
from cfjoin import xp

pytestmark = pytest.mark.gpu


def say(*a):
    print("CASE", *a, flush=True)


def sorted_side(rng, n, kmax):
    k = np.sort(rng.integers(0, kmax, n, dtype=np.uint64))
    i = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    return xp.asarray(k), xp.asarray(i)


def join_alloc(ctx, Dk, Did, Ik, Iid):
    od, oi, n = ctypes.c_void_p(), ctypes.c_void_p(), L.c_u64()
    L.check(L.load().cf_join_sorted(ctx.handle, ctypes.c_void_p(Dk.ptr), ctypes.c_void_p(Did.ptr), Dk.size,
                                    ctypes.c_void_p(Ik.ptr), ctypes.c_void_p(Iid.ptr), Ik.size, ctypes.byref(od),
                                    ctypes.byref(oi), ctypes.byref(n)), "cf_join_sorted")
    a = np.empty(n.value, np.uint64)
    b = np.empty(n.value, np.uint64)
    if n.value:
        ctx.d2h(a, od.value)
        ctx.d2h(b, oi.value)
        ctx.free(od.value)
        ctx.free(oi.value)
    return a, b


def join_noalloc(ctx, Dk, Did, Ik, Iid, out_d, out_i, cap):
    n = L.c_u64()
    rc = L.load().cf_join_sorted_noalloc(ctx.handle, ctypes.c_void_p(Dk.ptr), ctypes.c_void_p(Did.ptr), Dk.size,
                                         ctypes.c_void_p(Ik.ptr), ctypes.c_void_p(Iid.ptr), Ik.size,
                                         ctypes.c_void_p(out_d.ptr), ctypes.c_void_p(out_i.ptr), cap, ctypes.byref(n))
    return rc, n.value


def test_noalloc_join_equals_allocating_join():
    ctx = L.default_context()
    rng = np.random.default_rng(20261204)
    cap = 1 << 22
    out_d, out_i = xp.empty(cap), xp.empty(cap)
    # the largest I side allowed, then a much smaller one; keys drawn from a small range so runs repeat
    for nD, nI, kmax in ((300_000, L.CF_JOIN_NOALLOC_MAX_I, 1 << 21), (5_000, 777, 4_000)):
        say(f"no-alloc join nD={nD} nI={nI} keys < {kmax}")
        Dk, Did = sorted_side(rng, nD, kmax)
        Ik, Iid = sorted_side(rng, nI, kmax)
        ref_d, ref_i = join_alloc(ctx, Dk, Did, Ik, Iid)
        ctx.synchronize()
        before = ctx.pool_stats()
        rc, n = join_noalloc(ctx, Dk, Did, Ik, Iid, out_d, out_i, cap)
        assert rc == 0, L.load().cf_last_error()
        assert ctx.pool_stats() == before, "the no-allocation join touched the context pool"
        assert n == len(ref_d) > 0
        assert np.array_equal(out_d[:n].get(), ref_d) and np.array_equal(out_i[:n].get(), ref_i)
        # a capacity one short of the pair count: rejected, count reported, outputs untouched
        sentinel = np.full(4, 0xA5A5A5A5A5A5A5A5, np.uint64)
        od2, oi2 = xp.asarray(np.resize(sentinel, n)), xp.asarray(np.resize(sentinel, n))
        rc, n2 = join_noalloc(ctx, Dk, Did, Ik, Iid, od2, oi2, n - 1)
        assert rc == -1 and n2 == n
        assert np.all(od2.get() == sentinel[0]) and np.all(oi2.get() == sentinel[0])


def test_noalloc_join_rejects_oversized_i_side():
    ctx = L.default_context()
    d = xp.asarray(np.arange(4, dtype=np.uint64))
    big = xp.empty(L.CF_JOIN_NOALLOC_MAX_I + 1)
    out = xp.empty(8)
    rc, _ = join_noalloc(ctx, d, d, big, big, out, out, 8)
    assert rc == -1 and b"CF_JOIN_NOALLOC_MAX_I" in L.load().cf_last_error()
