"""GPU parity: the gfx950 library against the FP64 oracle.

In Thomas mode every kernel evaluates the reference's operations in the
reference's order (FP contraction off, IEEE division and square root), so
the bar is bit-for-bit equality of every field and scalar.  The partitioned
(DIRECT) mode changes the rounding of the tridiagonal solves only; its bar is
a relative difference of at most 1e-10 per field.
"""
import os

import numpy as np
import pytest

import lica_boiling1d_amd as L
from oracle import Oracle
from parity_util import exact_mismatches, max_rel, nearly_singular_poisson, thomas_longdouble, ulp_sensitivity

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def test_deck_case_200_steps_bit_exact():
    """The shipped deck (N=256, r=2), 200 steps, checked every 50 steps."""
    case = L.deck_case()
    _say("deck case N=256 r=2, 200 steps")
    b, o = L.Batch([case]), Oracle(case)
    assert o.init_status == 0
    assert exact_mismatches(b, 0, o) == []
    for _ in range(4):
        b.run(50)
        assert o.step(50) == 0
        assert exact_mismatches(b, 0, o) == []
    s = b.scalars()
    assert s["STEP"] == 200 and s["STATUS"] == 0
    assert s["PLACE"] > case.z0   # the vapour film grows


@pytest.mark.parametrize("n,r,steps", [(64, 1, 40), (100, 2, 40), (512, 3, 20), (1024, 4, 20)])
def test_sizes_bit_exact(n, r, steps):
    _say(f"size case n={n} r={r} steps={steps}")
    case = L.deck_case(n, r)
    b, o = L.Batch([case]), Oracle(case)
    b.run(steps)
    assert o.step(steps) == 0
    assert exact_mismatches(b, 0, o) == []


def test_single_large_case_bit_exact():
    """N=8192, r=4 with the seeded film perturbation, 10 steps."""
    case = L.single_large_case(seed=1)
    _say("single large case N=8192 r=4, 10 steps")
    b, o = L.Batch([case]), Oracle(case)
    b.run(10)
    assert o.step(10) == 0
    assert exact_mismatches(b, 0, o) == []


def test_large_case_1000_steps_bit_exact():
    """The single large case (N=8192, r=4, 10 K, 0.2 mm film with a seeded
    1 % perturbation) over its full 1000 steps in the reference's Thomas
    mode: every field, the band and the scalars equal the oracle's bit for
    bit at steps 250, 500 and 1000."""
    case = L.single_large_case(seed=1)
    b, o = L.Batch([case]), Oracle(case)
    done = 0
    for upto in (250, 500, 1000):
        _say(f"single large case N=8192 r=4, steps {done + 1}..{upto}")
        b.run(upto - done)
        assert o.step(upto - done) == 0
        done = upto
        assert exact_mismatches(b, 0, o) == [], upto


def test_heterogeneous_batch_bit_exact():
    """12 cases of different N, superheat and film thickness in one batch."""
    cases = L.batched_cases(12, seed=7, nmin=128, nmax=1024)
    _say("heterogeneous batch of 12, 15 steps")
    b = L.Batch(cases)
    b.run(15)
    for i, c in enumerate(cases):
        o = Oracle(c)
        assert o.step(15) == 0
        assert exact_mismatches(b, i, o) == [], f"case {i} n={c.n}"


def test_graph_replay_equals_stream_launches():
    cases = L.batched_cases(6, seed=3, nmin=128, nmax=512)
    _say("graph vs stream, 6 cases, 3 x 4 steps")
    a, g = L.Batch(cases), L.Batch(cases)
    g.use_graph(True)
    for _ in range(3):
        a.run(4)
        g.run(4)
        for i in range(len(cases)):
            for f in ("W", "P", "T", "PSIF", "PHI"):
                assert np.array_equal(a.field(f, i), g.field(f, i))
            assert np.array_equal(a.scalar_array(i), g.scalar_array(i))


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_PROJECTED])
def test_two_lanes_equal_one(mode):
    """The cases run as 2, 3, 4 or 6 sets (8 asked for, capped at the 6 cases)
    on as many streams (stream launches, graph replay and timing mode) give
    the single sequence's fields bit for bit."""
    cases = L.batched_cases(6, seed=5, nmin=128, nmax=512)
    _say(f"2-6 lanes vs one, 6 cases, mode {mode}, 2 x 4 steps")
    runs = []
    for lanes, graph, timing in ((1, False, False), (2, False, False), (2, True, False), (2, False, True),
                                 (3, False, True), (4, True, False), (4, False, False), (8, False, False)):
        b = L.Batch(cases)
        b.set_poisson(mode, 1e-10, 100)
        b.set_lanes(lanes)
        b.use_graph(graph)
        b.set_timing(timing)
        runs.append(b)
    for _ in range(2):
        for b in runs:
            b.run(4)
        for i in range(len(cases)):
            for f in ("W", "P", "T", "PSIF", "PHI"):
                ref = runs[0].field(f, i)
                for b in runs[1:]:
                    assert np.array_equal(ref, b.field(f, i)), (f, i)
            for b in runs[1:]:
                assert np.array_equal(runs[0].scalar_array(i), b.scalar_array(i)), i
    for b in runs:
        b.close()


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_DIRECT, L.POISSON_MIXED])
def test_block_size_does_not_change_results(mode):
    """Every reduction of the step is a maximum or ordered by the case, and
    the partitioned solves' chunks are set by the row count: 64, 256 and
    1024 threads per workgroup give the same bits in every mode."""
    case = L.deck_case(300, 2)
    _say(f"block sizes 64/256/1024, mode {mode}, 10 steps")
    ref = None
    for t in (64, 256, 1024):
        b = L.Batch([case])
        b.set_poisson(mode, 1e-10, 100)
        b.set_block(t)
        b.run(10)
        cur = [b.field(f) for f in ("W", "P", "T", "PSIF", "PHI")]
        if ref is None:
            ref = cur
        else:
            assert all(np.array_equal(x, y) for x, y in zip(ref, cur))


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_MIXED, L.POISSON_PROJECTED])
def test_binned_lanes_and_autotune_do_not_change_results(mode):
    """Four lanes as size bins of equal work, each with the launch shapes the
    autotuner picked by timing trial steps (threads per workgroup of every
    free kernel class, multigrid LDS), give the one-sequence default run's
    fields, scalars and per-case status bit for bit, and the autotuner
    leaves the batch in the state it found it in."""
    cases = L.batched_cases(48, seed=7, nmin=128, nmax=2048)
    _say(f"binned lanes + autotune vs one sequence, 48 cases, mode {mode}, 2 + 4 steps")
    ref, b = L.Batch(cases), L.Batch(cases)
    for x in (ref, b):
        x.set_poisson(mode, 1e-10, 100)
    ref.set_lanes(1)
    b.set_lanes(4)
    b.set_lane_split(True)
    for x in (ref, b):
        x.run_all(2)
    tab = b.autotune(1)
    _say("  " + str(tab))
    assert sum(t["cases"] for t in tab) == len(cases)
    assert [t["nmax"] for t in tab] == sorted((t["nmax"] for t in tab), reverse=True)
    for t in tab:
        assert all(64 <= v <= 1024 for v in t["threads"].values())
    for x in (ref, b):
        x.run_all(4)
    st_r, m_r = ref.progress()
    st_b, m_b = b.progress()
    assert np.array_equal(st_r, st_b) and np.array_equal(m_r, m_b)
    for i in range(len(cases)):
        for f in ("W", "P", "T", "PSIF", "PHI"):
            assert np.array_equal(ref.field(f, i), b.field(f, i), equal_nan=True), (f, i)
        assert np.array_equal(ref.scalar_array(i), b.scalar_array(i), equal_nan=True), i
    ref.close()
    b.close()


@pytest.mark.parametrize("which,steps", [("deck", 50), ("large", 10)])
def test_direct_mode_within_tolerance(which, steps):
    """Partitioned tridiagonal solves with double-double refinement.

    Bar: every field within 1e-10 relative of the oracle, or, for fields the
    problem itself leaves to rounding (W and P are rounding noise of a nearly
    motionless column at these times), within 10x the oracle's own change
    under a one-ulp perturbation of its input; the interface position within
    1e-10 relative."""
    case = L.deck_case() if which == "deck" else L.single_large_case(1)
    _say(f"direct (partitioned) tridiagonal mode, {which} case, {steps} steps")
    b, o = L.Batch([case]), Oracle(case)
    b.set_poisson(L.POISSON_DIRECT)
    b.run(steps)
    assert o.step(steps) == 0
    fields = ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI")
    rel = max_rel(b, 0, o, fields)
    sens = ulp_sensitivity(case, steps, fields)
    for f in fields:
        assert rel[f] <= max(1e-10, 10.0 * sens[f]), (f, rel[f], sens[f])
    zo = o.scalar_array()[3]
    assert abs(b.scalars()["PLACE"] - zo) <= 1e-10 * zo


def test_reference_call_sequence_matches_step():
    """determine_dt / rhsnlhs / convrge1 / transport as MAIN calls them."""
    _say("single-case call sequence from the deck files")
    lin, gin = os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in")
    ctx = L.Context(lin, gin)
    o = Oracle(L.deck_case())
    for _ in range(3):
        ctx.determine_dt()
        for msub in (1, 2, 3):
            w, p, zn, cn = ctx.rhsnlhs(msub)
            ctx.convrge1()
            ctx.transport()
            zn2, cn2 = ctx.grid_coupling(0)
        ctx.end_step()
    assert o.step(3) == 0
    assert np.array_equal(w, o.field("W")) and np.array_equal(p, o.field("P"))
    assert np.array_equal(zn2, o.field("PSI_ZN")) and np.array_equal(cn2, o.field("PSI_CN"))
    assert exact_mismatches(ctx, 0, o) == []


@pytest.mark.parametrize("mode,n", [(L.POISSON_THOMAS, 700), (L.POISSON_DIRECT, 700),
                                    (L.POISSON_DIRECT, 2048), (L.POISSON_DIRECT, 3000)])
def test_batched_poisson_systems(mode, n):
    """lica_poisson against a dense solve of the same POISSON matrices
    (2048 rows fill the register chunks of 256 threads exactly; 3000 rows
    exceed them and take the serial elimination)."""
    _say(f"batched poisson mode {mode} n={n}")
    rng = np.random.default_rng(11)
    nsys = 37
    a = -rng.uniform(0.5, 1.5, (nsys, n))
    c = -rng.uniform(0.5, 1.5, (nsys, n))
    a[:, 0] = 0.0
    b = -(a + c) - rng.uniform(0.01, 0.1, (nsys, n))
    g = rng.standard_normal((nsys, n))
    p = L.poisson(a, b, c, g, mode=mode)
    for s in range(nsys):
        A = np.diag(b[s]) + np.diag(a[s, 1:], -1) + np.diag(c[s, :-1], 1)
        x = np.linalg.solve(A, g[s])
        assert np.max(np.abs(p[s] - x)) <= 1e-10 * np.max(np.abs(x))


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_PROJECTED])
def test_checkpoint_restart_bit_exact(tmp_path, mode):
    """Binary checkpoint of the device state: save, load, continue ==
    uninterrupted (in the projected mode with its solution history)."""
    cases = L.batched_cases(4, seed=9, nmin=128, nmax=600)
    _say(f"checkpoint/restart, 4 cases, mode {mode}")
    a = L.Batch(cases)
    a.set_poisson(mode, 1e-10, 100)
    a.run(6)
    a.save(tmp_path / "ck.bin")
    a.run(6)
    b = L.Batch.load(tmp_path / "ck.bin")
    b.run(6)
    for i in range(len(cases)):
        for f in ("W", "P", "T", "PSIF", "PHI", "ALPHI", "MFLVS"):
            assert np.array_equal(a.field(f, i), b.field(f, i)), (i, f)
        assert np.array_equal(a.scalar_array(i), b.scalar_array(i))


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED, L.POISSON_PROJECTED])
@pytest.mark.parametrize("which,steps", [("deck", 60), ("large", 6)])
def test_multigrid_modes(mode, which, steps):
    """V-cycles (FP64, or FP32 corrections in FP64 defect correction, from
    zero or from the history projection) until the backward error and the
    last correction are both within 1e-10.  The direct mode's bar: every
    field within 1e-10 relative of the oracle or 10x the oracle's own change
    under a one-ulp perturbation of its input, the interface within 1e-10."""
    case = L.deck_case() if which == "deck" else L.single_large_case(1)
    _say(f"multigrid mode {mode}, {which} case, {steps} steps")
    b, o = L.Batch([case]), Oracle(case)
    b.set_poisson(mode, 1e-10, 100)
    b.run(steps)
    assert o.step(steps) == 0
    s = b.scalars()
    assert s["POISSON_SOLVES"] == 3 * steps
    assert 0 < s["POISSON_RESID"] <= 1e-10
    assert s["POISSON_CYCLES"] / s["POISSON_SOLVES"] <= (2.5 if mode == L.POISSON_MG64 else 6)
    fields = ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI")
    rel = max_rel(b, 0, o, fields)
    sens = ulp_sensitivity(case, steps, fields)
    _say("  " + " ".join(f"{f}={rel[f]:.1e}/{sens[f]:.1e}" for f in fields))
    for f in fields:
        assert rel[f] <= max(1e-10, 10.0 * sens[f]), (f, rel[f], sens[f])
    zo = o.scalar_array()[3]
    assert abs(s["PLACE"] - zo) <= 1e-10 * zo


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED])
def test_multigrid_modes_equal_logic_check_build(mode):
    """The same kernel source compiled for the CPU (work items in order, the
    logic-check build the CPU tests pin against the oracle) gives the GPU's
    fields and scalars bit for bit in the multigrid modes: the partitioned
    energy and momentum solves (rows in registers on the GPU, in arrays on
    the CPU) and the V-cycles (FP64, or FP32 in FP64 defect correction) do
    the same operations in the same order."""
    import emul_build
    from lica_boiling1d_amd import _abi
    if not os.path.exists(emul_build.LIB):
        pytest.fail("the logic-check build is missing: run __graft_entry__.build() first")
    emul = _abi.load(emul_build.LIB)
    cases = [L.deck_case(n, 2) for n in (100, 257, 600)]
    _say(f"multigrid mode {mode}: GPU vs logic-check build, n = 100, 257, 600, 12 steps")
    g, e = L.Batch(cases), L.Batch(cases, lib=emul)
    for b in (g, e):
        b.set_poisson(mode, 1e-10, 100)
        b.run(12)
    for i in range(len(cases)):
        for f in ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI"):
            assert np.array_equal(g.field(f, i), e.field(f, i)), (i, f)
        assert np.array_equal(g.scalar_array(i), e.scalar_array(i)), i
    g.close()
    e.close()


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED])
def test_batched_poisson_multigrid(mode):
    """lica_poisson by V-cycles on the POISSON systems of the deck case (the
    last substep of steps 1..8): backward error <= 1e-10 and the solution
    within 1e-12 of TRDIAG3's (whose own error on these rows is about 2e-15)."""
    _say(f"batched poisson multigrid mode {mode}")
    o = Oracle(L.deck_case())
    rows = []
    for _ in range(8):
        assert o.step(1) == 0
        rows.append(o.last_poisson()[:4])
    a, b, c, g = (np.array([r[j] for r in rows]) for j in range(4))
    p = L.poisson(a, b, c, g, mode=mode, rtol=1e-10, max_cycles=100)
    x = L.poisson(a, b, c, g, mode=L.POISSON_THOMAS)
    for s in range(len(rows)):
        r = g[s] - (b[s] * p[s] + np.r_[0.0, a[s, 1:] * p[s, :-1]] + np.r_[c[s, :-1] * p[s, 1:], 0.0])
        anorm = np.max(np.abs(a[s]) + np.abs(b[s]) + np.abs(c[s]))
        assert np.max(np.abs(r)) <= 1e-10 * (anorm * np.max(np.abs(p[s])) + np.max(np.abs(g[s])))
        assert np.max(np.abs(p[s] - x[s])) <= 1e-12 * np.max(np.abs(x[s]))


@pytest.mark.parametrize("mode,tol", [(L.POISSON_MG64, 1e-14), (L.POISSON_MIXED, 1e-11)])
def test_poisson_multigrid_nearly_singular_rows_gpu(mode, tol):
    """The batch's hardest POISSON rows (b = -(a + c) rounded, a 1605x jump
    of 1/rho, the Dirichlet row on top): the V-cycle solution within tol of
    an extended-precision elimination, on the GPU.  The reference's own
    elimination (the Thomas mode, TRDIAG3's order in FP64) is off by about
    1e-8 on the same rows, its pivots losing the row sums; that error, not
    the multigrid's, is what separates the multigrid runs' P from the Thomas
    run's on such cases."""
    _say(f"nearly singular POISSON rows n=3000, mode {mode}")
    a, b, c, g = nearly_singular_poisson()
    x = thomas_longdouble(a, b, c, g)
    scale = np.max(np.abs(x))
    p = L.poisson(a, b, c, g, mode=mode, rtol=1e-10, max_cycles=100)[0]
    assert np.max(np.abs(p - x)) <= tol * scale
    t = L.poisson(a, b, c, g, mode=L.POISSON_THOMAS)[0]
    assert np.max(np.abs(t - x)) >= 1e-10 * scale


def test_reference_seams_bit_exact():
    """MAIN's loop through the seams with the reference's own argument lists
    (arrays shaped (0:2,0:2,0:N3) in Fortran order, copied in at entry and
    out at exit), 3 steps of the deck: the oracle's fields and scalars bit
    for bit, the x/y arrays left zero."""
    from parity_util import check_reference_main
    _say("reference-signature seams, deck case, 3 steps")
    seams = L.ReferenceSeams(os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"))
    assert check_reference_main(seams, Oracle(L.deck_case()), 3, L.ref_array) == []
    seams.close()


@pytest.mark.parametrize("mode,tol", [(L.POISSON_THOMAS, 0.0), (L.POISSON_DIRECT, 1e-12), (L.POISSON_MG64, 1e-12),
                                      (L.POISSON_MIXED, 1e-11)])
def test_reference_poisson_seam(mode, tol):
    """POISSON(P,DIVGSUM,DENF_XI,DENF_YI,DENF_ZI,DENF_Y,DENF_C) on the inputs
    of the oracle's last substep: P(1,1,1:N) the oracle's TRDIAG3 solution
    (bit for bit in Thomas mode), P(1,1,0) and P(1,1,N3) its end copies."""
    from oracle.binding import trdiag3
    _say(f"POISSON seam, mode {mode}")
    o = Oracle(L.deck_case())
    assert o.step(3) == 0
    dzi, dg = o.last_poisson_inputs()
    a, b, c, g, _ = o.last_poisson()
    x = trdiag3(a, b, c, g)
    seams = L.ReferenceSeams(os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"))
    seams.set_poisson(mode, 1e-10, 100)
    n = seams.n
    P, DZI = L.ref_array(n), L.ref_array(n)
    DZI[:, 1, 1] = dzi
    seams.POISSON(P, dg, None, None, DZI, None, None)
    col = P[:, 1, 1]
    assert np.max(np.abs(col[1:n + 1] - x)) <= tol * np.max(np.abs(x))
    assert col[0] == col[1] and col[n + 1] == col[n]
    seams.close()


def test_fld_writer_and_restart(tmp_path):
    """The field file the GPU writes after 6 steps of the deck case holds the
    oracle's W, P, T and band as WRITEFIELD prints them (ES20.12: every
    value parses back to the oracle's value printed that way).  Restarting
    from it (IREAD=1, ITRACKING=1: PREFLD, LVSINIT from the band read,
    GRID_COUPLING(1)) as the second case of a three-case batch gives the
    oracle's restart fields, and 6 steps later its fields and scalars, bit
    for bit, while the batch's other cases run on as in a batch that was
    never touched."""
    from oracle.binding import read_fld
    case = L.deck_case()
    _say("field file of the deck case after 6 steps: writer vs oracle, restart as case 1 of 3")
    g, o = L.Batch([case]), Oracle(case)
    g.run(6)
    assert o.step(6) == 0
    path = tmp_path / "fld000006"
    g.write_fld(path, 0, 6)
    d = read_fld(path)
    es = lambda v: float("%20.12E" % v)
    assert d["n3"] == case.n + 1 and d["u0"]
    assert np.array_equal(d["w"], [es(x) for x in o.field("W")[1:]])
    assert np.array_equal(d["p"], [es(x) for x in o.field("P")[1:case.n + 1]])
    assert np.array_equal(d["t"], [es(x) for x in o.field("T")])
    bandf, alf = o.field("BAND"), o.field("ALPHI")
    pos, ai = 0, 0
    for planes in d["band"]:
        cnt = int(bandf[pos])
        assert [k for k, _ in planes] == [int(x) for x in bandf[pos + 1:pos + 1 + cnt]]
        assert [a for _, a in planes] == [es(x) for x in alf[ai:ai + cnt]]
        pos += 1 + cnt
        ai += cnt
    g.close()
    others = [L.deck_case(300, 2), case, L.deck_case(200, 2)]
    b, ref = L.Batch(others), L.Batch(others)
    for x in (b, ref):
        x.run(2)
    assert b.read_fld(path, 1, ireset=False, idtold=True) == 0
    o2 = Oracle(case, init=False)
    assert o2.init_fld(path, ireset=False, idtold=True) == 0
    fields = ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI")
    for f in fields:
        assert np.array_equal(b.field(f, 1), o2.field(f)), f
    for x in (b, ref):
        x.run(6)
    assert o2.step(6) == 0
    for f in fields:
        assert np.array_equal(b.field(f, 1), o2.field(f)), f
        for i in (0, 2):
            assert np.array_equal(b.field(f, i), ref.field(f, i)), (f, i)
    assert np.array_equal(b.scalar_array(1)[:9], o2.scalar_array()[:9])
    b.close()
    ref.close()
