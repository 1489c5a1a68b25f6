"""cfjoin: MI355X-native exchange-MITM residue join (HIP kernels behind a C ABI).

Modules:
  instance  exact instance recipe (pool, bounds, strata)
  verify    host re-check of hits (exact completion + Korselt)
  lib       ctypes binding of libcfjoin.so (the HIP kernels)
  xp        the small device-array module the drop-in modules return
  pipeline  plain and batched solves
"""
__all__ = ["instance", "verify"]
