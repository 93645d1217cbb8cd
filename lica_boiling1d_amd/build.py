"""Build recipes for the gfx950 library of the boiling path.

``build_library()`` compiles ``csrc/`` with hipcc for gfx950 into
``lica_boiling1d_amd/liblica_mi355x.so`` (in-tree, next to this file).
FP contraction is off so that every multiply and add rounds on its own,
which is what the bit-for-bit comparison against the serial FP64 order
relies on; packed FP32 instructions are disabled for the whole library.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblica_mi355x.so")
SOURCES = ["lica_kernels.hip", "lica_host.cpp", "lica_net.hip"]
HEADERS = ["lica_types.h", "lica_par.h", "lica_launch.h", "lica_kern_common.h",
           "lica_kern_band.h", "lica_kern_bandreg.h", "lica_kern_tgrad.h", "lica_kern_flow.h", "lica_kern_mg.h"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.sep not in p or os.path.exists(p)):
            return p
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    """Compile the HIP library for gfx950; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
