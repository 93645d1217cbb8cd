"""GPU parity on a fixed sample of the seeded 4096-case batch (the batched
FP64, mixed-precision and learned warm-start configurations), 25 steps as the
benchmark runs them (5 warmup + 20 timed).

Sample (indices into batched_cases(4096), seed 20261202): the four largest
cases (N = 4090..4094); cases that reach one of the reference's STOP branches
in the oracle at steps 1 to 8 (CAL_PSIF_MTP's error branch,
two_phase.f:2905, and non-finite fields); the two cases whose P the
multigrid runs of the whole batch move farthest from the Thomas run's
(2974 and 888: POISSON systems of condition number 4-5e10, where TRDIAG3's
own error is 1.5e-8 to 5e-8); and ordinary cases of N = 222..3893.

Bars:
* Thomas mode: every case stops at the same step with the same status as the
  oracle, and every field and scalar of the running cases equals the
  oracle's bit for bit.
* Direct (partitioned, double-double refined) and multigrid modes: every
  case stops at the same step with the same status as the oracle, except a
  case whose stop the Poisson solve's rounding decides -- one where the
  direct mode, accurate to rounding, and the oracle's TRDIAG3 disagree (case
  3398 of the batch: TRDIAG3's pressure sends it into CAL_PSIF_MTP's error
  branch at step 1, a solution accurate to rounding does not); the interface
  of the running cases within 1e-10 relative of the oracle; W, P, T, PHI
  within 1e-10 relative or 10x the direct mode's own difference from the
  oracle.
"""
import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd.learned import Net, collect_pairs
from oracle import Oracle
from parity_util import exact_mismatches, max_rel

pytestmark = pytest.mark.gpu

SAMPLE = [2135, 680, 2967, 2176, 3398, 299, 2, 57, 8, 48, 80, 169, 732, 95, 144, 190, 103, 3363, 1948, 1480, 1535,
          2039, 2276, 31, 1419, 3163, 1918, 782, 113, 3523, 3610, 2852, 2974, 888]
STEPS = 25
FIELDS = ("W", "P", "T", "PHI")


def _say(msg):
    print(msg, flush=True)


def _cases():
    allc = L.batched_cases(4096)
    return [allc[i] for i in SAMPLE]


@pytest.fixture(scope="module")
def oracle_runs():
    """(status, steps entered -- the one a case stops in included, as the
    library counts them --, fields, scalars) of every sample case in the
    oracle."""
    out = []
    for c in _cases():
        o = Oracle(c)
        st, done = 0, 0
        for _ in range(STEPS):
            st = o.step(1)
            done += 1
            if st:
                break
        out.append((st, done, {f: o.field(f) for f in FIELDS}, o.scalar_array(), o))
    return out


def _run(mode, net=None):
    b = L.Batch(_cases())
    if net is not None:
        b.set_net(net)
    b.set_poisson(mode, 1e-10, 100)
    b.run_all(STEPS)
    st, steps = b.progress()
    return b, st, steps


def _rel(b, i, o):
    out = {f: float(np.max(np.abs(b.field(f, i) - o[2][f])) / max(np.max(np.abs(o[2][f])), 1e-300)) for f in FIELDS}
    zo = o[3][3]
    out["PLACE"] = abs(b.scalars(i)["PLACE"] - zo) / abs(zo)
    return out


@pytest.fixture(scope="module")
def direct_run(oracle_runs):
    b, st, steps = _run(L.POISSON_DIRECT)
    rel = [(_rel(b, i, o) if st[i] == 0 and o[0] == 0 else None) for i, o in enumerate(oracle_runs)]
    odd = {SAMPLE[i] for i, o in enumerate(oracle_runs) if (st[i], steps[i]) != (o[0], o[1])}
    b.close()
    return st, steps, rel, odd


def test_thomas_sample_bit_exact(oracle_runs):
    _say(f"batch sample of {len(SAMPLE)} cases, Thomas mode, {STEPS} steps")
    b, st, steps = _run(L.POISSON_THOMAS)
    for i, o in enumerate(oracle_runs):
        assert (st[i], steps[i]) == (o[0], o[1]), (SAMPLE[i], st[i], steps[i], o[0], o[1])
        if o[0] == 0:
            assert exact_mismatches(b, i, o[4]) == [], SAMPLE[i]
    b.close()


def test_direct_sample(oracle_runs, direct_run):
    st, steps, rel, odd = direct_run
    _say(f"batch sample, direct mode: cases whose stop the rounding of the Poisson solve decides {sorted(odd)}")
    assert odd <= {3398}, odd
    for i, o in enumerate(oracle_runs):
        if rel[i] is not None:
            assert rel[i]["PLACE"] <= 1e-10, (SAMPLE[i], rel[i])


def _check_mode(mode, oracle_runs, direct_run, net=None):
    dst, dsteps, drel, odd = direct_run
    b, st, steps = _run(mode, net)
    cyc = []
    for i, o in enumerate(oracle_runs):
        if SAMPLE[i] in odd:
            continue
        assert (st[i], steps[i]) == (o[0], o[1]), (SAMPLE[i], st[i], steps[i], o[0], o[1])
        if o[0] != 0:
            continue
        s = b.scalars(i)
        assert 0 < s["POISSON_RESID"] <= 1e-10, (SAMPLE[i], s["POISSON_RESID"])
        cyc.append(s["POISSON_CYCLES"] / s["POISSON_SOLVES"])
        r = _rel(b, i, o)
        assert r["PLACE"] <= 1e-10, (SAMPLE[i], r)
        for f in FIELDS:
            assert r[f] <= max(1e-10, 10.0 * drel[i][f]), (SAMPLE[i], f, r[f], drel[i][f])
    b.close()
    return float(np.mean(cyc))


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED, L.POISSON_PROJECTED])
def test_multigrid_sample(oracle_runs, direct_run, mode):
    _say(f"batch sample, multigrid mode {mode}, {STEPS} steps")
    cyc = _check_mode(mode, oracle_runs, direct_run)
    _say(f"  V-cycles per solve {cyc:.2f}")


def test_learned_sample(oracle_runs, direct_run):
    """The network trained on pairs of a disjoint batch (seed 20261203)."""
    _say("batch sample, learned mode: train on 64 cases of a disjoint batch, then 25 steps")
    net = Net(seed=0)
    tc = L.batched_cases(64, seed=20261203)
    assert collect_pairs(net, tc, steps=8, windows_per_case=4, window=256, warmup=4) > 0
    net.train(300, batch=32, lr=1e-3)
    cyc = _check_mode(L.POISSON_LEARNED, oracle_runs, direct_run, net)
    _say(f"  V-cycles per solve {cyc:.2f}")
    net.close()
