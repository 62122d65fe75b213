"""Builds the HIP libraries for gfx950 with hipcc: whir_amd/libwhir_amd.so
(KoalaBear) and whir_amd/libwhir_amd_bb.so (the same sources for BabyBear).

All kernels live in one translation unit (csrc/whir_amd.hip) because they share
the __constant__ Poseidon2 tables. Packed-FP32 instruction selection is
disabled for every HIP compile.

The compile keeps its intermediate files in a temporary directory so that the
gfx950 assembly of the hashing kernels can be read back: their VALU
instruction counts (straight-line code, so per lane per launch) go to
whir_amd/kernel_stats.json, which bench.py uses for the VALU roofline.
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwhir_amd.so")
# the same sources built for the reference's second field (BabyBear)
LIB_BB = os.path.join(HERE, "libwhir_amd_bb.so")
STATS = os.path.join(HERE, "kernel_stats.json")
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

# kernels whose per-lane VALU instruction count bench.py needs (mangled-name
# prefixes: kernel name and template arguments, any parameter list)
COUNTED = {
    "leaves16": "_ZN2kb20merkle_leaves_kernelILi16EEEv",
    "leaves64": "_ZN2kb20merkle_leaves_kernelILi64EEEv",
    "layer": "_ZN2kb19merkle_layer_kernelE",
}


def count_valu(asm, prefix):
    m = re.search(r"^(" + re.escape(prefix) + r"\w*):", asm, re.M)
    if not m:
        return None
    i = m.start()
    j = asm.find(".Lfunc_end", i)
    n = 0
    for line in asm[i:j].splitlines():
        t = line.strip()
        if t.startswith("v_"):
            n += 1
    return n


def build(verbose=False, extra_flags=()):
    """Both libraries: KoalaBear (with the VALU counts bench.py reads) and BabyBear."""
    build_field("koalabear", verbose, extra_flags)
    build_field("babybear", verbose, extra_flags)
    return LIB


def build_field(field, verbose=False, extra_flags=()):
    if field == "babybear":
        return _build_one(LIB_BB, ["-DWHIR_FIELD_BABYBEAR", *extra_flags], verbose, stats=False)
    return _build_one(LIB, list(extra_flags), verbose, stats=True)


def _build_one(lib_path, extra_flags, verbose, stats):
    src = os.path.join(CSRC, "whir_amd.hip")
    with tempfile.TemporaryDirectory(prefix="whir_build_") as tmp:
        out = os.path.join(tmp, "libwhir_amd.so")
        cmd = [HIPCC, *HIP_FLAGS, *extra_flags, "-save-temps", "-I", CSRC, src, "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd, cwd=tmp)
        asm_files = [f for f in os.listdir(tmp) if f.endswith(".s") and ARCH in f]
        counts = {"arch": ARCH}
        if stats and asm_files:
            with open(os.path.join(tmp, asm_files[0])) as fh:
                asm = fh.read()
            for key, sym in COUNTED.items():
                counts[key + "_valu_per_lane"] = count_valu(asm, sym)
        shutil.copyfile(out, lib_path + ".tmp")
        os.replace(lib_path + ".tmp", lib_path)
    if stats:
        with open(STATS, "w") as fh:
            json.dump(counts, fh, indent=1)
    return lib_path


if __name__ == "__main__":
    build(verbose=True, extra_flags=sys.argv[1:])
