"""Builds the CPU logic-check library of the kernel source (tests only).

The kernel headers in lica_boiling1d_amd/csrc compile as plain C++ when
LICA_EMUL is defined: a workgroup then runs its work items in order.  The
result, tests/_emul/liblica_emul.so, exposes the same C ABI and lets the
CPU tests compare the kernels' logic with the oracle bit for bit.  It is
never loaded by the package.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "lica_boiling1d_amd", "csrc")
OUT = os.path.join(HERE, "_emul")
LIB = os.path.join(OUT, "liblica_emul.so")


def build(force=False):
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in ("lica_kernels.hip", "lica_host.cpp")]
    deps = srcs + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(p) for p in deps):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-DLICA_EMUL", "-ffp-contract=off", "-fPIC", "-shared", "-pthread",
           "-x", "c++", *srcs, "-o", tmp]
    subprocess.run(cmd, check=True, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
