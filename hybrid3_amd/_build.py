"""Builds libhybrid3.so (HIP kernels + C-ABI) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libhybrid3.so")
MAIN = os.path.join(HERE, "hybrid3")  # `hybrid3 N NB [IB]`, the reference's main.c
SOURCES = ["hq_factor.hip", "hq_generic.hip", "hq_homegrown.hip", "hq_runtime.cpp", "hq_harness.cpp"]
ARCH = os.environ.get("HQ_OFFLOAD_ARCH", "gfx950")

# Packed FP32 ops are switched off for every device compile (the hardware
# forms of v_pk_* FP32 arithmetic are not used by this library).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                # keep MFMA accumulators in ordinary VGPRs: the strip of the
                # updated tile is both an MFMA accumulator and an MFMA operand,
                # and the AGPR form copies it between register files each block
                "-mllvm", "-amdgpu-mfma-vgpr-form=1",
                # every atomic read-modify-write here is issued by one lane; the
                # optimizer's wave-wide rewrite would wait for the look-ahead
                # claim at once instead of phases later
                "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]


def hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.sep not in p or os.path.exists(p)):
            return p
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
           "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
           *DEVICE_FLAGS, "-I", os.path.join(HERE, "..", "include"),
           "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libhybrid3.so failed")
    os.replace(tmp, LIB)
    # the reference's driver program on the library (plain C, gcc)
    cmd = ["gcc", "-O2", "-std=c11", "-D_POSIX_C_SOURCE=199309L", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(CSRC, "hq_main.c"), "-o", MAIN + ".tmp", "-L", HERE, "-lhybrid3",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of the hybrid3 driver failed")
    os.replace(MAIN + ".tmp", MAIN)
    if verbose:
        print("built", LIB, "and", MAIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
