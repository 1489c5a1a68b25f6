"""ctypes binding of libcfjoin.so (include/cfjoin.h).

The library must have been built (cfjoin/build.py); there is no fallback of any kind: a missing
library raises ImportError at first use, a failed call raises. Status mapping follows the
reference's error behaviour: a violated precondition raises AssertionError (as the asserts of
gpu/mitm_gpu2.py:83-88 and gpu/igen_gpu.py:103-105 do), a device allocation failure raises
MemoryError, a HIP runtime failure raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(HERE, "libcfjoin.so")

c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
P_u64 = ctypes.POINTER(ctypes.c_uint64)

CF_ID_COLEX, CF_ID_PACK_ASC, CF_ID_PACK_DESC, CF_ID_MASK = 0, 1, 2, 3
CF_SOLVE_PARTITION, CF_SOLVE_SORTMERGE, CF_SOLVE_PARTITION_DGLOBAL = 0, 1, 2
CF_JOIN_NOALLOC_MAX_I = 1 << 20

_lib = None
_lock = threading.RLock()


class cf_instance(ctypes.Structure):
    _fields_ = [("M", c_u64), ("r", ctypes.c_int32), ("nB", ctypes.c_int32), ("B", c_void_p),
                ("nI", ctypes.c_int32), ("hmax", ctypes.c_int32), ("INS", c_void_p), ("P0modM", c_u64),
                ("icap_lo", c_u64), ("icap_hi", c_u64), ("dmin_lo", c_u64), ("dmin_hi", c_u64),
                ("split_idx", ctypes.c_int32), ("i_id_mode", ctypes.c_int32), ("d_id_mode", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("M_hi", c_u64), ("P0modM_hi", c_u64)]


class cf_stats(ctypes.Structure):
    _fields_ = [("nD", c_u64), ("nI", c_u64), ("npairs", c_u64), ("peak_bytes", c_u64),
                ("ms_total", ctypes.c_double), ("ms_gen_d", ctypes.c_double), ("ms_sort_d", ctypes.c_double),
                ("ms_gen_i", ctypes.c_double), ("ms_sort_i", ctypes.c_double), ("ms_join", ctypes.c_double),
                ("nD_parents", c_u64), ("nI_parents", c_u64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class cf_tuning(ctypes.Structure):
    _fields_ = [("slice_kib", ctypes.c_int32), ("d_per_bucket", ctypes.c_int32), ("seg_fill_pct", ctypes.c_int32),
                ("probe_blocks", ctypes.c_int32)]


class cf_served(ctypes.Structure):
    _fields_ = [("tag", c_u64), ("status", ctypes.c_int32), ("batch_size", ctypes.c_int32), ("nD", c_u64),
                ("nI", c_u64), ("npairs", c_u64), ("pairs_d", ctypes.POINTER(c_u64)),
                ("pairs_i", ctypes.POINTER(c_u64)), ("t_submit_us", ctypes.c_double),
                ("t_start_us", ctypes.c_double), ("t_done_us", ctypes.c_double), ("error", ctypes.c_char * 160)]


# name -> (restype, argtypes); every name here is declared in include/cfjoin.h
SIGNATURES = {
    "cf_version": (c_int, []),
    "cf_last_error": (ctypes.c_char_p, []),
    "cf_ctx_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "cf_ctx_destroy": (c_int, [c_void_p]),
    "cf_alloc": (c_int, [c_void_p, c_u64, ctypes.POINTER(c_void_p)]),
    "cf_free": (c_int, [c_void_p, c_void_p]),
    "cf_pool_release": (c_int, [c_void_p]),
    "cf_pool_stats": (c_int, [c_void_p, P_u64, P_u64]),
    "cf_mem_info": (c_int, [c_void_p, P_u64, P_u64]),
    "cf_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_u64]),
    "cf_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_u64]),
    "cf_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_void_p, c_u64]),
    "cf_synchronize": (c_int, [c_void_p]),
    "cf_prof_enable": (c_int, [c_void_p, ctypes.c_char_p]),
    "cf_prof_disable": (c_int, [c_void_p]),
    "cf_prof_reset": (c_int, [c_void_p]),
    "cf_prof_read": (c_int, [c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), P_u64]),
    "cf_prof_dump": (c_int, [c_void_p, ctypes.c_char_p, c_u64]),
    "cf_gen_subsets": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_u64, c_u64, c_u64, c_u64, c_void_p, c_void_p]),
    "cf_sort_pairs": (c_int, [c_void_p, c_void_p, c_void_p, c_u64]),
    "cf_sort_pairs_into": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_u64, c_u64]),
    "cf_join_sorted": (c_int, [c_void_p, c_void_p, c_void_p, c_u64, c_void_p, c_void_p, c_u64,
                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), P_u64]),
    "cf_join_sorted_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_u64, c_void_p, c_void_p, c_u64, c_void_p,
                                       c_void_p, c_u64, P_u64]),
    "cf_plan_create_i": (c_int, [c_void_p, c_void_p, c_int, c_int, c_u64, c_u64, c_u64, c_u64, c_int, c_int, c_int,
                                 ctypes.POINTER(c_void_p)]),
    "cf_plan_create_d": (c_int, [c_void_p, c_void_p, c_int, c_int, c_u64, c_u64, c_u64, c_int,
                                 ctypes.POINTER(c_void_p)]),
    "cf_plan_total": (c_int, [c_void_p, P_u64]),
    "cf_plan_first_counts": (c_int, [c_void_p, c_void_p]),
    "cf_plan_write": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, P_u64]),
    "cf_plan_destroy": (c_int, [c_void_p]),
    "cf_solve": (c_int, [c_void_p, ctypes.POINTER(cf_instance), c_int, c_u64, ctypes.POINTER(c_void_p),
                         ctypes.POINTER(c_void_p), P_u64, ctypes.POINTER(cf_stats)]),
    "cf_solve_batch": (c_int, [c_void_p, ctypes.POINTER(cf_instance), c_int, ctypes.POINTER(c_void_p),
                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), P_u64, P_u64,
                               ctypes.POINTER(cf_stats)]),
    "cf_batch_create": (c_int, [c_void_p, ctypes.POINTER(cf_instance), c_int, ctypes.POINTER(c_void_p)]),
    "cf_batch_solve": (c_int, [c_void_p, c_u64, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                               ctypes.POINTER(c_void_p), P_u64, P_u64, ctypes.POINTER(cf_stats)]),
    "cf_batch_destroy": (c_int, [c_void_p]),
    "cf_batch_capture": (c_int, [c_void_p, c_u64]),
    "cf_batch_set_tuning": (c_int, [c_void_p, ctypes.POINTER(cf_tuning)]),
    "cf_batch_footprint": (c_int, [c_void_p, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "cf_peak_valu": (c_int, [c_void_p, c_int, c_u64, P_u64, ctypes.POINTER(ctypes.c_double)]),
    "cf_server_create": (c_int, [c_int, c_int, c_int, c_u64, ctypes.POINTER(c_void_p)]),
    "cf_server_submit": (c_int, [c_void_p, c_void_p, c_u64]),
    "cf_server_next": (c_int, [c_void_p, c_int, ctypes.POINTER(cf_served)]),
    "cf_server_release": (c_int, [c_void_p, c_u64]),
    "cf_server_pending": (c_int, [c_void_p, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "cf_server_destroy": (c_int, [c_void_p]),
}


def load():
    """Load libcfjoin.so and declare every exported signature. Raises ImportError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(SO_PATH):
                raise ImportError(f"libcfjoin.so is not built ({SO_PATH}); run cfjoin/build.py")
            L = ctypes.CDLL(SO_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == 0:
        return
    msg = (load().cf_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == -1:
        raise AssertionError(text)
    if rc == -2:
        raise MemoryError(text)
    raise RuntimeError(f"{text} (status {rc})")


class Context:
    """One device + one HIP stream + a caching device pool (cf_ctx)."""

    def __init__(self, device: int = 0):
        L = load()
        h = c_void_p()
        check(L.cf_ctx_create(device, ctypes.byref(h)), "cf_ctx_create")
        self.handle = h
        self.device = device

    def alloc(self, nbytes: int) -> int:
        p = c_void_p()
        check(load().cf_alloc(self.handle, max(int(nbytes), 1), ctypes.byref(p)), "cf_alloc")
        return p.value

    def free(self, ptr: int) -> None:
        if ptr and self.handle:
            load().cf_free(self.handle, c_void_p(ptr))

    def release_pool(self) -> None:
        check(load().cf_pool_release(self.handle), "cf_pool_release")

    def pool_stats(self) -> Tuple[int, int]:
        a, b = c_u64(), c_u64()
        check(load().cf_pool_stats(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def mem_info(self) -> Tuple[int, int]:
        f, t = c_u64(), c_u64()
        check(load().cf_mem_info(self.handle, ctypes.byref(f), ctypes.byref(t)), "cf_mem_info")
        return f.value, t.value

    def h2d(self, dst: int, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host)
        check(load().cf_memcpy_h2d(self.handle, c_void_p(dst), host.ctypes.data_as(c_void_p), host.nbytes), "h2d")

    def d2h(self, host: np.ndarray, src: int) -> None:
        check(load().cf_memcpy_d2h(self.handle, host.ctypes.data_as(c_void_p), c_void_p(src), host.nbytes), "d2h")

    def d2d(self, dst: int, src: int, nbytes: int) -> None:
        check(load().cf_memcpy_d2d(self.handle, c_void_p(dst), c_void_p(src), nbytes), "d2d")

    def synchronize(self) -> None:
        check(load().cf_synchronize(self.handle), "cf_synchronize")

    def prof_enable(self, kernel: Optional[str] = None) -> None:
        check(load().cf_prof_enable(self.handle, kernel.encode() if kernel else None), "cf_prof_enable")

    def prof_disable(self) -> None:
        check(load().cf_prof_disable(self.handle), "cf_prof_disable")

    def prof_reset(self) -> None:
        check(load().cf_prof_reset(self.handle), "cf_prof_reset")

    def prof_dump(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        check(load().cf_prof_dump(self.handle, buf, len(buf)), "cf_prof_dump")
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.split()
            out[name] = (float(ms), int(n))
        return out

    def prof_read(self, kernel: str) -> Tuple[float, int]:
        ms, n = ctypes.c_double(), c_u64()
        check(load().cf_prof_read(self.handle, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)), "cf_prof_read")
        return ms.value, n.value


_default: Optional[Context] = None


def default_context() -> Context:
    global _default
    if _default is None:
        with _lock:
            if _default is None:
                _default = Context(int(os.environ.get("CF_DEVICE", "0")))
    return _default


def split128(v: int) -> Tuple[int, int]:
    if v < 0 or v >= 1 << 128:
        raise AssertionError("value needs more than 128 bits")
    return v & ((1 << 64) - 1), v >> 64


def u64_array(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


class Plan:
    """A filtered enumeration plan (cf_plan): the tree is expanded on creation."""

    def __init__(self, handle, ctx: Context, n: int):
        self.handle = handle
        self.ctx = ctx
        self.n = n

    @classmethod
    def insertion(cls, INS, r: int, M: int, P0modM: int, Icap: int, split_idx: int, hmax: int,
                  id_mode: int = CF_ID_COLEX, ctx: Optional[Context] = None) -> "Plan":
        ctx = ctx or default_context()
        ins = u64_array(INS)
        lo, hi = split128(max(Icap, 0))
        h = c_void_p()
        check(load().cf_plan_create_i(ctx.handle, ins.ctypes.data_as(c_void_p), len(ins), r, M, P0modM, lo, hi,
                                      int(split_idx), int(hmax), id_mode, ctypes.byref(h)), "cf_plan_create_i")
        return cls(h, ctx, len(ins))

    @classmethod
    def deletion(cls, B, r: int, M: int, Dmin: int, id_mode: int = CF_ID_PACK_DESC,
                 ctx: Optional[Context] = None) -> "Plan":
        ctx = ctx or default_context()
        b = u64_array(B)
        lo, hi = split128(Dmin)
        h = c_void_p()
        check(load().cf_plan_create_d(ctx.handle, b.ctypes.data_as(c_void_p), len(b), r, M, lo, hi, id_mode,
                                      ctypes.byref(h)), "cf_plan_create_d")
        return cls(h, ctx, len(b))

    def total(self) -> int:
        t = c_u64()
        check(load().cf_plan_total(self.handle, ctypes.byref(t)), "cf_plan_total")
        return t.value

    def first_counts(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        check(load().cf_plan_first_counts(self.handle, out.ctypes.data_as(c_void_p)), "cf_plan_first_counts")
        return out

    def write(self, first_lo: int, first_hi: int, keys_ptr: int, ids_ptr: int) -> int:
        w = c_u64()
        check(load().cf_plan_write(self.handle, first_lo, first_hi, c_void_p(keys_ptr), c_void_p(ids_ptr),
                                   ctypes.byref(w)), "cf_plan_write")
        return w.value

    def close(self) -> None:
        if self.handle:
            load().cf_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
