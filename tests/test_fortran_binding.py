"""The Fortran binding INTEGRATION.md shows, compiled and run: its
ISO_C_BINDING module (taken from INTEGRATION.md as it stands) and a driver
that runs MAIN's loop (1d.f:196-295) on the reference's own arrays
(tests/fortran/main_steps.f90), built with amdflang and linked to the CPU
logic-check build of the kernel source (tests/_emul, the same C ABI as the
GPU library).  Three steps of the shipped deck give the FP64 oracle's
fields and scalars bit for bit.  CPU only."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import emul_build
import lica_boiling1d_amd as L
from conftest import GOLDEN, ROOT
from oracle import Oracle


def _flang():
    for p in (shutil.which("amdflang"), "/opt/rocm/bin/amdflang", "/opt/rocm/llvm/bin/amdflang"):
        if p and os.path.exists(p):
            return p
    return None


def module_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```fortran\n(.*?)```", text, flags=re.S)
    mods = [b for b in blocks if b.lstrip().lower().startswith("module lica_mi355x")]
    assert len(mods) == 1, "INTEGRATION.md must hold exactly one lica_mi355x module block"
    return mods[0]


def test_fortran_driver_matches_oracle(tmp_path):
    flang = _flang()
    if flang is None:
        pytest.skip("amdflang is not installed")
    lib = emul_build.build()
    libdir = os.path.dirname(lib)
    mod = tmp_path / "lica_mi355x_mod.f90"
    mod.write_text(module_source())
    drv = os.path.join(ROOT, "tests", "fortran", "main_steps.f90")
    exe = tmp_path / "main_steps"
    subprocess.run([flang, "-O1", str(mod), drv, "-o", str(exe), f"-L{libdir}", "-llica_emul",
                    f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    case = L.deck_case()
    steps = 3
    out = tmp_path / "out.txt"
    subprocess.run([str(exe), os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"), str(case.n + 1),
                    str(steps), str(out)], check=True, cwd=tmp_path, timeout=300)
    rows = np.loadtxt(out)
    assert rows.shape == (case.n + 3, 5)
    o = Oracle(case)
    assert o.step(steps) == 0
    for j, f in enumerate(("W", "P", "PSI_ZN", "PSI_CN", "T")):
        assert np.array_equal(rows[:-1, j], o.field(f)), f
    s = o.scalar_array()
    assert tuple(rows[-1]) == (s[4], s[5], s[1], s[6], s[0])    # DVMAX, CFLM, DT, VOL_TOT_ORI, TIME
