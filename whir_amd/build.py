"""Builds the HIP library whir_amd/libwhir_amd.so for gfx950 with hipcc.

All kernels live in one translation unit (csrc/whir_amd.hip) because they share
the __constant__ Poseidon2 tables. Packed-FP32 instruction selection is
disabled for every HIP compile.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwhir_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WHIR_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, extra_flags=()):
    src = os.path.join(CSRC, "whir_amd.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, *extra_flags, "-I", CSRC, src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra_flags=sys.argv[1:])
