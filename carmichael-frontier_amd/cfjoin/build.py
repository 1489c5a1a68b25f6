"""Compile libcfjoin.so (the HIP kernels and the C ABI) in-tree for gfx950.

    python -m cfjoin.build            (from carmichael-frontier_amd/)

The library is written next to this file; lib.py loads it from there and nowhere else.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
SO = os.path.join(HERE, "libcfjoin.so")
SOURCES = ["k_gen.hip", "k_sort.hip", "k_join.hip", "k_part.hip", "capi.hip", "solve.hip", "serve.hip", "k_peak.hip", "k_wide.hip"]
ARCH = os.environ.get("CF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def command(out: str = SO) -> list:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
            # packed fp32 forms are kept out of the code object (the kernels are integer-only)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(os.path.dirname(PKG_ROOT), "include"),
            "-o", out] + srcs


def build(verbose: bool = False) -> str:
    tmp = SO + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    errs = [l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs[-60:]))
    os.replace(tmp, SO)
    return SO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
