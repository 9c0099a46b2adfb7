"""Builds libqrhip.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python qrusty_amd/build.py   (or __graft_entry__.build()); it does not
import the package, which would load the library being rebuilt.  Each source
is compiled to its own object in qrusty_amd/_build/ (in parallel), then the
objects are linked.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqrhip.so")
SOURCES = ["qr_api.cpp", "qr_csr.hip", "qr_apply.hip", "qr_batch.hip", "qr_group.hip",
           "qr_vec.hip", "qr_spadd.hip", "qr_filter.hip", "qr_io.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no fused multiply-add unless written as fma(): TwoSum and the bit-exact
    # CSR values depend on it
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
LINK = ["--offload-arch=gfx950", "-fPIC", "-shared"]

# zlib for gzip-compressed CSR files
LIBS = ["-lz"]


def _jobs() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(len(SOURCES), n, 8))


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        return src, obj, cmd, res

    with ThreadPoolExecutor(_jobs()) as pool:
        results = list(pool.map(compile_one, SOURCES))
    failed = False
    for src, _, cmd, res in results:
        if verbose:
            print(" ".join(cmd), flush=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            failed = True
        elif verbose:
            # hipcc's host pass does not know the packed-FP32 feature: harmless
            err = "\n".join(l for l in res.stderr.splitlines()
                            if "packed-fp32-ops" not in l)
            if err.strip():
                sys.stderr.write(err + "\n")
    if failed:
        raise RuntimeError("hipcc failed building libqrhip.so")
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + LINK + ["-o", tmp] + [obj for _, obj, _, _ in results] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libqrhip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
