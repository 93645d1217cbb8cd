"""Host-side logic of bench.py on synthetic run records (CPU only): the
comparison of a run with the Thomas run (stops case by case, the fields of
the cases alive in both) and the whole-step roofline."""
import importlib.util
import os

import numpy as np
import pytest

import lica_boiling1d_amd as L
from conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _record(bench, status, steps, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    n = len(status)
    fields = {f: [rng.uniform(1.0, 2.0, 9) * scale for _ in range(n)] for f in bench.FIELD_CMP}
    return {"status": np.array(status), "steps": np.array(steps), "place": np.full(n, 0.06),
            "qwall": np.full(n, 2.0), "fields": fields}


def test_versus_thomas_counts_stops_and_compares_live_fields(bench):
    """A case stopped here but alive in the Thomas run counts as such and is
    left out of the field comparison; a case stopped in both at another step
    counts as another stop; the fields' per-case relative differences give
    the largest and the median over the cases alive in both."""
    ref = _record(bench, [0, 0, 3, 0, 5], [10, 10, 4, 10, 7], seed=1)
    res = _record(bench, [0, 2, 3, 0, 5], [10, 6, 4, 10, 8], seed=1)
    rel = {0: 1e-9, 3: 4e-9}
    for f in bench.FIELD_CMP:
        for i, e in rel.items():
            res["fields"][f][i] = ref["fields"][f][i] * (1.0 + e)
        res["fields"][f][1] = ref["fields"][f][1] * 10.0   # stopped here: not compared
    res["qwall"] = ref["qwall"].copy()
    res["qwall"][3] *= 1.0 + 2e-12
    out = bench.versus_thomas(res, ref)
    assert out["cases_with_other_stop_than_thomas"] == 2
    assert out["those_cases"] == [1, 4]
    assert out["alive_in_thomas_but_stopped_here"] == 1
    cmp = out["field_rel_diff_vs_thomas"]
    assert cmp["cases"] == 2
    for f in bench.FIELD_CMP:
        assert cmp["max"][f] == pytest.approx(4e-9, rel=1e-6)
        assert cmp["median"][f] == pytest.approx(2.5e-9, rel=1e-6)
    assert cmp["max"]["QWALL"] == pytest.approx(2e-12, rel=1e-3)
    assert cmp["max"]["PLACE"] == 0.0


def test_versus_thomas_without_fields(bench):
    """Runs recorded without fields are compared by stops and interface only."""
    ref = _record(bench, [0, 0], [3, 3], seed=2)
    res = _record(bench, [0, 0], [3, 3], seed=2)
    res["fields"] = None
    out = bench.versus_thomas(res, ref)
    assert "field_rel_diff_vs_thomas" not in out
    assert out["interface_rel_diff_max_over_cases_alive_in_both"] == 0.0


def test_step_roofline_sums_completed_steps(bench):
    """The whole-step figure is the per-kernel byte rule summed over the
    kernels of a step (DETERMINE_DT once, the substep's kernels three
    times), times the steps each case completed, over the window's time."""
    cases = L.batched_cases(8)
    done = np.array([5, 5, 0, 3, 5, 5, 1, 5])
    res = {"el": 0.25, "done_per_case": done}
    for mode in (L.POISSON_THOMAS, L.POISSON_MIXED, L.POISSON_LEARNED):
        r = bench.step_roofline(cases, mode, res)
        per = [bench.kernel_bytes(c, "dt", mode)
               + 3 * sum(bench.kernel_bytes(c, k, mode) for k in bench.STEP_KERNELS) for c in cases]
        tot = sum(p * int(d) for p, d in zip(per, done))
        assert r["algorithmic_bytes"] == tot
        assert r["achieved"] == pytest.approx(tot / 0.25 / 1e9)
        assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    # the warm-start base counts only in the warm-started modes
    c = cases[0]
    assert bench.kernel_bytes(c, "warm_start", L.POISSON_MIXED) == 0
    assert bench.kernel_bytes(c, "warm_start", L.POISSON_PROJECTED) > 0
    assert bench.kernel_bytes(c, "warm_start", L.POISSON_LEARNED) > 0
