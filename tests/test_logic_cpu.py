"""CPU logic check of the kernel source: the same HIP kernel bodies built
as plain C++ (tests/_emul, LICA_EMUL: a workgroup runs its work items in
order and aborts on a barrier in divergent code) against the oracle, bit
for bit in Thomas mode.  The GPU tests repeat these comparisons on the
device build."""
import numpy as np
import pytest

import emul_build
import lica_boiling1d_amd as L
import lica_boiling1d_amd._abi as abi
from oracle import Oracle
from parity_util import exact_mismatches, max_rel


@pytest.fixture(scope="module")
def emul():
    return abi.load(emul_build.build())


def test_deck_case_bit_exact(emul):
    case = L.deck_case()
    b, o = L.Batch([case], lib=emul), Oracle(case)
    assert exact_mismatches(b, 0, o) == []
    for _ in range(3):
        b.run(50)
        assert o.step(50) == 0
        assert exact_mismatches(b, 0, o) == []


@pytest.mark.parametrize("n,r,steps", [(64, 1, 30), (100, 3, 20), (257, 2, 25)])
def test_sizes_bit_exact(emul, n, r, steps):
    case = L.deck_case(n, r)
    b, o = L.Batch([case], lib=emul), Oracle(case)
    b.run(steps)
    assert o.step(steps) == 0
    assert exact_mismatches(b, 0, o) == []


def test_heterogeneous_batch_bit_exact(emul):
    cases = L.batched_cases(4, seed=11, nmin=128, nmax=400)
    b = L.Batch(cases, lib=emul)
    b.run(12)
    for i, c in enumerate(cases):
        o = Oracle(c)
        assert o.step(12) == 0
        assert exact_mismatches(b, i, o) == [], i


def test_direct_mode_close_to_oracle(emul):
    case = L.deck_case()
    b, o = L.Batch([case], lib=emul), Oracle(case)
    b.set_poisson(L.POISSON_DIRECT)
    b.run(30)
    assert o.step(30) == 0
    rel = max_rel(b, 0, o, ("T", "PSI_CN", "PSI_ZN", "PSIF", "PHI"))
    assert all(v <= 1e-10 for v in rel.values()), rel


def test_stopped_case_keeps_status_others_continue(emul):
    """A case that reaches a reference STOP (here the CAL_PSIF_MTP error
    branch after a degenerate face box, as the oracle does too) keeps its
    status; the rest of the batch advances."""
    cases = L.batched_cases(4096)
    bad = cases[299]      # N=3273: stops at step 2 in the oracle as well
    good = L.deck_case()
    o = Oracle(bad)
    st_o = 0
    for s in range(3):
        st_o = o.step(1)
        if st_o:
            break
    b = L.Batch([good, bad], lib=emul)
    nbad = b.run_all(3)
    st, steps = b.progress()
    assert nbad == 1 and st[1] == st_o != 0 and st[0] == 0 and steps[0] == 3
