"""Whole-instance solves on the device.

solve(inst)        -- one instance per call ("plain"): every equal-residue pair of the instance.
                      The device half of gpu/mitm_gpu_mstar_driver.py:209-243 (DFILTER=1,
                      IGEN=gpu) in one C call; mode "partition" (default) partitions I leaves by
                      key range and probes the D buckets of their partition, mode "sortmerge"
                      materializes, sorts and joins I chunks as the reference does.
solve_batch(insts) -- many instances of any radii in one solve (one key space, one D bucketing,
                      one set of launches per radius).
decode(inst, pairs) / recheck(inst, pairs) -- host decode and exact re-check of the hits
               (gpu/mitm_gpu_mstar_driver.py:224-237).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import lib as L
from . import verify


@dataclass
class SolveResult:
    pairs_d: np.ndarray          # D ids (uint64)
    pairs_i: np.ndarray          # I ids (uint64)
    nD: int
    nI: int
    stats: dict = field(default_factory=dict)

    @property
    def keys_joined(self) -> int:
        return self.nD + self.nI

    def pair_array(self) -> np.ndarray:
        """Pairs as sorted (D id, I id) rows: the canonical form compared for bit-exactness."""
        a = np.stack([self.pairs_d, self.pairs_i], axis=1) if len(self.pairs_d) else np.zeros((0, 2), np.uint64)
        if len(a):
            a = a[np.lexsort((a[:, 1], a[:, 0]))]
        return a


def id_modes(inst):
    i_mode = L.CF_ID_COLEX if inst.rank_ids else L.CF_ID_PACK_ASC
    d_mode = L.CF_ID_MASK if inst.r > 10 else L.CF_ID_PACK_DESC
    return i_mode, d_mode


def make_cinstance(inst, keep: list) -> L.cf_instance:
    B = L.u64_array(inst.B0)
    INS = L.u64_array(inst.INS)
    keep += [B, INS]
    i_mode, d_mode = id_modes(inst)
    ilo, ihi = L.split128(inst.Icap)
    dlo, dhi = L.split128(inst.Dmin)
    mlo, mhi = L.split128(inst.M)
    plo, phi = L.split128(inst.P0 % inst.M)
    return L.cf_instance(M=mlo, r=inst.r, nB=len(B), B=B.ctypes.data_as(ctypes.c_void_p), nI=len(INS),
                         hmax=int(inst.hmax), INS=INS.ctypes.data_as(ctypes.c_void_p), P0modM=plo,
                         icap_lo=ilo, icap_hi=ihi, dmin_lo=dlo, dmin_hi=dhi, split_idx=int(inst.split_idx),
                         i_id_mode=i_mode, d_id_mode=d_mode, pad=0, M_hi=mhi, P0modM_hi=phi)


MODES = {"partition": L.CF_SOLVE_PARTITION, "sortmerge": L.CF_SOLVE_SORTMERGE,
         "partition_dglobal": L.CF_SOLVE_PARTITION_DGLOBAL}


def solve(inst, ctx: Optional[L.Context] = None, max_chunk: int = 0, mode: str = "partition") -> SolveResult:
    ctx = ctx or L.default_context()
    keep: List = []
    ci = make_cinstance(inst, keep)
    od, oi = ctypes.c_void_p(), ctypes.c_void_p()
    n = L.c_u64()
    st = L.cf_stats()
    L.check(L.load().cf_solve(ctx.handle, ctypes.byref(ci), MODES[mode], max_chunk, ctypes.byref(od),
                              ctypes.byref(oi), ctypes.byref(n), ctypes.byref(st)), "cf_solve")
    pd = np.empty(n.value, np.uint64)
    pi = np.empty(n.value, np.uint64)
    if n.value:
        ctx.d2h(pd, od.value)
        ctx.d2h(pi, oi.value)
        ctx.free(od.value)
        ctx.free(oi.value)
    return SolveResult(pd, pi, int(st.nD), int(st.nI), st.as_dict())


class Batch:
    """Instances prepared on the device (cf_batch_create): descriptors and exact bound tables are
    uploaded once; solve() is device work only and can be repeated."""

    def __init__(self, insts, ctx: Optional[L.Context] = None):
        self.ctx = ctx or L.default_context()
        self.insts = list(insts)
        keep: List = []
        n = len(self.insts)
        arr = (L.cf_instance * n)(*[make_cinstance(x, keep) for x in self.insts])
        h = ctypes.c_void_p()
        L.check(L.load().cf_batch_create(self.ctx.handle, arr, n, ctypes.byref(h)), "cf_batch_create")
        self.handle = h

    def set_tuning(self, slice_kib: int = 0, d_per_bucket: int = 0, seg_fill_pct: int = 0,
                   probe_blocks: int = 0) -> None:
        """Speed knobs of later solves (cf_batch_set_tuning); 0 keeps a default. Results never change."""
        t = L.cf_tuning(slice_kib, d_per_bucket, seg_fill_pct, probe_blocks)
        L.check(L.load().cf_batch_set_tuning(self.handle, ctypes.byref(t)), "cf_batch_set_tuning")
        self.tuning = dict(slice_kib=slice_kib, d_per_bucket=d_per_bucket, seg_fill_pct=seg_fill_pct,
                           probe_blocks=probe_blocks)

    def capture(self, max_chunk: int = 0) -> None:
        """Capture the batch's solve into a HIP graph (cf_batch_capture): later solve() calls with
        max_chunk 0 (or the captured one) are one graph launch each."""
        L.check(L.load().cf_batch_capture(self.handle, max_chunk), "cf_batch_capture")
        self.captured = True

    def solve(self, max_chunk: int = 0):
        """Returns (list of SolveResult in input order, stats dict)."""
        ctx = self.ctx
        n = len(self.insts)
        od, oi, oinst = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        npairs = L.c_u64()
        counts = np.zeros(2 * n, np.uint64)
        st = L.cf_stats()
        L.check(L.load().cf_batch_solve(self.handle, max_chunk, ctypes.byref(od), ctypes.byref(oi),
                                        ctypes.byref(oinst), ctypes.byref(npairs),
                                        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(st)),
                "cf_batch_solve")
        m = npairs.value
        pd = np.empty(m, np.uint64)
        pi = np.empty(m, np.uint64)
        pj = np.empty(m, np.uint64)
        if m:
            for host, dev in ((pd, od), (pi, oi), (pj, oinst)):
                ctx.d2h(host, dev.value)
                ctx.free(dev.value)
        order = np.argsort(pj, kind="stable")
        pd, pi, pj = pd[order], pi[order], pj[order]
        # each instance's pairs are a view of its own range of these freshly sorted arrays
        bounds = np.searchsorted(pj, np.arange(n + 1, dtype=np.uint64)).tolist()
        cl = counts.tolist()
        out = [SolveResult(pd[bounds[j]:bounds[j + 1]], pi[bounds[j]:bounds[j + 1]], cl[2 * j], cl[2 * j + 1])
               for j in range(n)]
        return out, st.as_dict()

    def close(self) -> None:
        if getattr(self, "handle", None):
            L.load().cf_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(insts, ctx: Optional[L.Context] = None):
    """Solve many instances together. Returns (list of SolveResult in input order, batch stats)."""
    b = Batch(insts, ctx)
    try:
        return b.solve()
    finally:
        b.close()


def decoders(inst):
    """(D id -> index tuple, I id -> index tuple) for the instance's id encodings."""
    if inst.r > 10:
        d_dec = verify.decode_bitmask
    else:
        d_dec = lambda v: verify.decode_packed(v, inst.r, 6)  # noqa: E731
    if inst.rank_ids:
        i_dec = lambda v: verify.unrank_colex(v, inst.r, len(inst.INS))  # noqa: E731
    else:
        i_dec = lambda v: verify.decode_packed(v, inst.r, 8)  # noqa: E731
    return d_dec, i_dec


def recheck(inst, res: SolveResult) -> List[dict]:
    """Exact host re-check of every pair (residue identity, integrality, Korselt)."""
    d_dec, i_dec = decoders(inst)
    return verify.recheck_pairs(inst, res.pairs_d, res.pairs_i, d_dec, i_dec)
