"""The C ABI: every entry point include/lica_mi355x.h declares is exported
by the gfx950 library; deck parsing and argument errors (host code only,
no GPU work); without a GPU the library refuses to run."""
import os
import re
import shutil
import subprocess

import pytest

import lica_boiling1d_amd as L
import lica_boiling1d_amd._abi as abi
from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "lica_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lica_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("lica_init", "lica_step", "lica_rhsnlhs", "lica_transport", "lica_grid_coupling",
                 "lica_determine_dt", "lica_convrge1", "lica_poisson", "lica_batch_create", "lica_batch_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("library not built")
    names = declared_functions()
    nm = shutil.which("nm")
    if nm:
        out = subprocess.run([nm, "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
        exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
        missing = [n for n in names if n not in exported]
    else:
        lib = abi.load()
        missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_parse_shipped_deck():
    d = L.parse_deck(os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"))
    assert (d.n, d.r, d.idtopt) == (256, 2, 1)
    assert (d.re, d.visr, d.denr) == (1.42e-2, 22.22, 1605.36)
    assert (d.surf_j, d.fr, d.prm, d.scr, d.tcr) == (7.36e-8, 9.7, 1.0231, 2.077, 27.16)
    assert (d.tsat, d.ja, d.dt0, d.cflmax, d.zl) == (1.0, 0.009, 2.5e-6, 0.5, 1.0)
    assert d.as_dict() == L.deck_case().as_dict()


def test_parse_rejects_other_configurations(tmp_path):
    lines = open(os.path.join(GOLDEN, "lica.in")).read().splitlines()
    idx = [i for i, l in enumerate(lines) if l.startswith("ICH")][0] + 1
    lines[idx] = "0    1     1     1"          # IPZ=1: periodic z, not this column
    p = tmp_path / "lica.in"
    p.write_text("\n".join(lines) + "\n")
    d = L.CaseDesc()
    lib = abi.load()
    import ctypes
    rc = lib.lica_parse_deck(str(p).encode(), os.path.join(GOLDEN, "grid.in").encode(), ctypes.byref(d))
    assert rc == abi.UNSUPPORTED_DECK
    rc = lib.lica_parse_deck(b"/nonexistent/lica.in", os.path.join(GOLDEN, "grid.in").encode(), ctypes.byref(d))
    assert rc == abi.IO


def test_no_gpu_means_no_run():
    """The product path has no CPU implementation: without a device the
    library reports NO_DEVICE instead of computing anything."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(L.LicaError) as e:
        L.Batch([L.deck_case()])
    assert e.value.status == abi.NO_DEVICE


def test_bench_measures_one_gpu_only():
    """`bench.py --gpus N` with N > 1 exits with a message instead of
    running: the batch's cases would shard by case, one GPU is measured."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
