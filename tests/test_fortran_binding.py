"""The Fortran binding compiled and run: its ISO_C_BINDING module
(include/lica_mi355x.f90) and a driver that runs MAIN's loop (1d.f:196-295)
on the reference's own arrays (tests/fortran/main_steps.f90), built with
amdflang and linked to the CPU logic-check build of the kernel source
(tests/_emul, the same C ABI as the GPU library).  Three steps of the
shipped deck give the FP64 oracle's fields and scalars bit for bit.  The
module block the repository's markdown pages show is the file itself.  CPU
only."""
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


MODULE = os.path.join(ROOT, "include", "lica_mi355x.f90")


def module_source():
    with open(MODULE) as f:
        return f.read()


def test_shown_module_is_the_shipped_file():
    """Every lica_mi355x module block in the top-level markdown pages equals
    include/lica_mi355x.f90, and at least one page shows it."""
    shown = []
    for name in sorted(os.listdir(ROOT)):
        if name.endswith(".md"):
            with open(os.path.join(ROOT, name)) as f:
                blocks = re.findall(r"```fortran\n(.*?)```", f.read(), flags=re.S)
            shown += [b for b in blocks if b.lstrip().lower().startswith("module lica_mi355x")]
    assert shown
    assert all(b == module_source() for b in shown)


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
