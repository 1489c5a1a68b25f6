"""A small HIP-backed array module: exactly the array surface the reference drivers and
validators use on the join path, implemented over libcfjoin's device pool.

Surface (with the reference call sites it serves):
  ndarray .size .shape .dtype .nbytes len() .get() .item() slicing [a:b]   (gpu/mitm_gpu_mstar_driver.py:218,222;
                                                                           gpu/mitm_gpu2.py:124)
  asarray(numpy) / asnumpy(x)                  H2D / D2H (gpu/mitm_gpu_mstar_driver.py:112,224-225,254)
  empty(n, dtype) / zeros(n, dtype)            device allocation (gpu/mitm_gpu2.py:90)
  concatenate(list)                            device-to-device copies (gpu/igen_gpu.py:152)
  sort(x)                                      device radix sort (gpu/validate_mitm_gpu2.py:66)
  get_default_memory_pool().free_all_blocks()  pool release (gpu/mitm_gpu_mstar_driver.py:114,238)
  cuda.Device().mem_info / .synchronize()      (gpu/mitm_gpu_mstar_driver.py:75)
Elementwise comparison (==, all) exists only for validation scripts; it compares on the host.
Only 8-byte unsigned/signed integers are stored on the device.
"""
from __future__ import annotations

from typing import Iterable, Optional

import numpy as np

from . import lib as _lib

uint64 = np.uint64
int64 = np.int64
_ALLOWED = (np.dtype(np.uint64), np.dtype(np.int64))


class _Buffer:
    """One pool allocation; freed back to the pool when the last view goes away."""
    __slots__ = ("ptr", "nbytes", "ctx")

    def __init__(self, nbytes: int, ctx: _lib.Context, ptr: Optional[int] = None):
        self.ctx = ctx
        self.nbytes = nbytes
        self.ptr = ptr if ptr is not None else ctx.alloc(nbytes)

    def __del__(self):
        try:
            self.ctx.free(self.ptr)
        except Exception:
            pass


class ndarray:
    """1-D device array of 8-byte integers."""
    __slots__ = ("_buf", "_off", "size", "dtype", "key_bound", "__weakref__")

    def __init__(self, size: int, dtype=np.uint64, _buf: Optional[_Buffer] = None, _off: int = 0,
                 ctx: Optional[_lib.Context] = None):
        dt = np.dtype(dtype)
        if dt not in _ALLOWED:
            raise TypeError(f"device arrays hold 8-byte integers, not {dt}")
        self.dtype = dt
        self.size = int(size)
        if _buf is None:
            _buf = _Buffer(max(self.size * 8, 8), ctx or _lib.default_context())
        self._buf = _buf
        self._off = _off
        # every value is below this bound when the library's generator that wrote them says so
        # (keys reduced mod M); None when unknown. Slices and concatenations do not carry it.
        self.key_bound = None

    @classmethod
    def _adopt(cls, ptr: int, size: int, dtype=np.uint64, ctx: Optional[_lib.Context] = None) -> "ndarray":
        """Take ownership of a pool allocation returned by the library."""
        ctx = ctx or _lib.default_context()
        return cls(size, dtype, _buf=_Buffer(max(size * 8, 8), ctx, ptr=ptr))

    @property
    def ptr(self) -> int:
        return self._buf.ptr + self._off

    @property
    def ctx(self) -> _lib.Context:
        return self._buf.ctx

    @property
    def shape(self):
        return (self.size,)

    @property
    def nbytes(self) -> int:
        return self.size * 8

    @property
    def ndim(self) -> int:
        return 1

    def __len__(self) -> int:
        return self.size

    def get(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        if self.size:
            self.ctx.d2h(out, self.ptr)
        return out

    def item(self):
        if self.size != 1:
            raise ValueError("item() needs a size-1 array")
        return self.get()[0].item()

    def copy(self) -> "ndarray":
        out = ndarray(self.size, self.dtype, ctx=self.ctx)
        if self.size:
            self.ctx.d2d(out.ptr, self.ptr, self.nbytes)
        out.key_bound = self.key_bound
        return out

    def astype(self, dtype, copy: bool = True) -> "ndarray":
        out = self.copy() if copy else self
        out.dtype = np.dtype(dtype)
        if out.dtype not in _ALLOWED:
            raise TypeError("device arrays hold 8-byte integers")
        return out

    def __getitem__(self, key):
        if isinstance(key, slice):
            start, stop, step = key.indices(self.size)
            if step != 1:
                raise IndexError("device arrays support contiguous slices only")
            n = max(0, stop - start)
            return ndarray(n, self.dtype, _buf=self._buf, _off=self._off + start * 8)
        if isinstance(key, (int, np.integer)):
            k = int(key)
            if k < 0:
                k += self.size
            if not 0 <= k < self.size:
                raise IndexError(key)
            return self[k:k + 1].get()[0]
        raise IndexError("device arrays support integer and contiguous slice indexing")

    def __eq__(self, other):  # validation convenience: compared on the host
        o = other.get() if isinstance(other, ndarray) else other
        return self.get() == o

    __hash__ = None

    def __repr__(self) -> str:
        return f"cfjoin.xp.ndarray(size={self.size}, dtype={self.dtype})"


def empty(n: int, dtype=np.uint64) -> ndarray:
    return ndarray(int(n), dtype)


def zeros(n: int, dtype=np.uint64) -> ndarray:
    return asarray(np.zeros(int(n), dtype=dtype))


def asarray(a, dtype=None) -> ndarray:
    if isinstance(a, ndarray):
        return a if dtype is None or np.dtype(dtype) == a.dtype else a.astype(dtype)
    h = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    if h.dtype not in _ALLOWED:
        h = h.astype(np.uint64)
    h = h.reshape(-1)
    out = ndarray(h.size, h.dtype)
    if h.size:
        out.ctx.h2d(out.ptr, h)
    return out


def asnumpy(a) -> np.ndarray:
    return a.get() if isinstance(a, ndarray) else np.asarray(a)


def concatenate(arrs: Iterable[ndarray]) -> ndarray:
    arrs = list(arrs)
    if not arrs:
        return empty(0)
    dt = arrs[0].dtype
    out = ndarray(sum(a.size for a in arrs), dt)
    at = 0
    for a in arrs:
        if a.size:
            out.ctx.d2d(out.ptr + at * 8, a.ptr, a.nbytes)
        at += a.size
    return out


def sort(a: ndarray) -> ndarray:
    """Sorted copy (stable device radix sort; uint64 order)."""
    k = a.copy()
    scratch = empty(a.size)
    if a.size > 1:
        _lib.check(_lib.load().cf_sort_pairs(k.ctx.handle, _lib.c_void_p(k.ptr), _lib.c_void_p(scratch.ptr), k.size),
                   "cf_sort_pairs")
    return k


def all(x) -> bool:  # noqa: A001 - mirrors the array-module name the validators call
    return bool(np.all(x.get() if isinstance(x, ndarray) else x))


class _Pool:
    def free_all_blocks(self) -> None:
        _lib.default_context().release_pool()

    def used_bytes(self) -> int:
        return _lib.default_context().pool_stats()[0]

    def total_bytes(self) -> int:
        live, cached = _lib.default_context().pool_stats()
        return live + cached


_pool = _Pool()


def get_default_memory_pool() -> _Pool:
    return _pool


class _Device:
    def __init__(self, device: int = 0):
        self.id = device

    @property
    def mem_info(self):
        return _lib.default_context().mem_info()

    def synchronize(self) -> None:
        _lib.default_context().synchronize()


class cuda:  # namespace mirroring the array-module layout the drivers call (cp.cuda.Device())
    Device = _Device
