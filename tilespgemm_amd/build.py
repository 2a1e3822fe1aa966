"""Build recipe for libtsg.so (gfx950) -- compiles every HIP source in csrc/
with hipcc and links the shared library in-tree, next to this file.

Flags: -O3, C++17, -ffp-contract=off (FP64 sums are either explicit FMAs or
separate multiply/add, never contracted behind the code's back), packed-FP32
instructions disabled for device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtsg.so")
CLI = os.path.join(HERE, "tsg_spgemm")
SOURCES = ["tsg_util.hip", "tsg_row.hip", "tsg_tile.hip", "tsg_batch.hip", "tsg_server.hip",
           "tsg_api.hip", "tsg_io.hip"]
ARCH = os.environ.get("TSG_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc(), *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile and link libtsg.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libtsg.so failed")
    os.replace(tmp, LIB)
    # command-line driver, linked against the library next to it
    cmd = [hipcc(), *flags(), os.path.join(CSRC, "tsg_cli.hip"), "-o", CLI + ".tmp",
           f"-L{HERE}", "-ltsg", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of tsg_spgemm failed")
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
