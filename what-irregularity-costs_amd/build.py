"""Build recipe for the gfx950 TSDF library.

Compiles every csrc/*.hip file with hipcc for gfx950 and links them into
libosn_tsdf_hip.so next to this file (in-tree, so the built library travels
with the source tree to the machine that runs it).

Flags:
  -ffp-contract=off   the band walk must round like the reference CPU arm
                      (x86-64 SSE, no FMA) for the allocated key set to be
                      bit-identical; intended fused multiply-adds are
                      written explicitly;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from forming packed FP32 ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libosn_tsdf_hip.so")
ARCH = os.environ.get("OSN_TSDF_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-value"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops", "unknown target feature")
    err = "\n".join(l for l in r.stderr.splitlines() if not any(n in l for n in noise))
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_harness(verbose)
    build_ceiling(verbose)
    return LIB


HARNESS_SRC = os.path.join(HERE, "harness", "osn_harness.cpp")
HARNESS = os.path.join(HERE, "osn_harness")


def build_harness(verbose: bool = False) -> str:
    """The timing harness: host C++ over the C ABI, linked to the library next
    to it (rpath $ORIGIN, so the pair travels together)."""
    tmp = HARNESS + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", HARNESS_SRC, "-o", tmp, f"-L{HERE}",
           "-losn_tsdf_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"harness build failed:\n{r.stderr}")
    os.replace(tmp, HARNESS)
    if verbose:
        print(f"built {HARNESS}")
    return HARNESS


CEILING_SRC = os.path.join(HERE, "harness", "rmw_ceiling.hip")
CEILING = os.path.join(HERE, "rmw_ceiling")


def build_ceiling(verbose: bool = False) -> str:
    """The memory-ceiling probe of the update's access pattern (standalone)."""
    tmp = CEILING + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-Wall", CEILING_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rmw_ceiling build failed:\n{r.stderr}")
    os.replace(tmp, CEILING)
    if verbose:
        print(f"built {CEILING}")
    return CEILING


if __name__ == "__main__":
    build(verbose=True)
