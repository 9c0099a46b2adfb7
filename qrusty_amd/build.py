"""Builds libqrhip.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python qrusty_amd/build.py   (or __graft_entry__.build()); it does not
import the package, which would load the library being rebuilt.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqrhip.so")
SOURCES = ["qr_api.cpp", "qr_csr.hip", "qr_apply.hip", "qr_batch.hip", "qr_vec.hip",
           "qr_spadd.hip", "qr_filter.hip", "qr_io.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no fused multiply-add unless written as fma(): TwoSum and the bit-exact
    # CSR values depend on it
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


# zlib for gzip-compressed CSR files
LIBS = ["-lz"]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libqrhip.so")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
