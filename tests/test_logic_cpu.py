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
from parity_util import exact_mismatches, max_rel, nearly_singular_poisson, thomas_longdouble


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


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_PROJECTED])
def test_checkpoint_restart_bit_exact(emul, tmp_path, mode):
    """save after 7 steps, load, 8 more == 15 uninterrupted steps (the
    projected warm start's solution history and mode travel with the file)"""
    cases = L.batched_cases(3, seed=5, nmin=128, nmax=300)
    a = L.Batch(cases, lib=emul)
    a.set_poisson(mode, 1e-10, 100)
    a.run(7)
    a.save(tmp_path / "ck.bin")
    a.run(8)
    b = L.Batch.load(tmp_path / "ck.bin", lib=emul)
    b.run(8)
    for i in range(len(cases)):
        for f in ("W", "P", "T", "PSIF", "PHI", "ALPHI", "MFLVS", "SFGAS"):
            assert np.array_equal(a.field(f, i), b.field(f, i)), (i, f)
        assert np.array_equal(a.scalar_array(i), b.scalar_array(i))


def test_writefield_file(emul, tmp_path):
    """The WRITEFIELD layout (1d.f:1121-1232): headers, U/V/W/P blocks, band
    list with r x r columns per plane, T block; values to 12 digits."""
    case = L.deck_case()
    b = L.Batch([case], lib=emul)
    b.run(5)
    path = tmp_path / "fld000005"
    b.write_fld(path, 0, 5)
    toks = path.read_text().split()
    n, r = int(case.n), int(case.r)
    N1, N2, N3 = 2, 2, n + 1
    assert [int(t) for t in toks[:3]] == [N1, N2, N3]
    assert int(toks[3]) == 5 and int(toks[4]) == 5
    assert float(toks[5]) == pytest.approx(b.scalars()["TIME"], rel=1e-11)
    pos = 5 + 2 + 4 + 1 + 1                      # TIME DT, IPS..PRA, QVOL_ORI
    pos += N1 * (N2 + 1) * (N3 + 1) + (N1 + 1) * N2 * (N3 + 1)
    w = np.array([float(t) for t in toks[pos:pos + (N1 + 1) * (N2 + 1) * N3]]).reshape(N3, N2 + 1, N1 + 1)
    assert np.allclose(w[:, 1, 1], b.field("W")[1:], rtol=1e-11, atol=1e-300)
    pos += (N1 + 1) * (N2 + 1) * N3
    p = np.array([float(t) for t in toks[pos:pos + n]])
    assert np.allclose(p, b.field("P")[1:n + 1], rtol=1e-11, atol=1e-300)
    pos += n
    assert [int(t) for t in toks[pos:pos + 3]] == [r + 1, r + 1, r * n + 1]
    pos += 3 + 6 + 2 + 1
    numa = [int(t) for t in toks[pos:pos + 11]]
    band = b.field("BAND")
    assert sum(numa) == r * r * (len(band) - 11)
    pos += 11 + 4 * sum(numa) + 1 + 1 + 1 + 1 + 1 + 1 + 3
    t = np.array([float(x) for x in toks[pos:]]).reshape(N3 + 1, N2 + 1, N1 + 1)
    assert np.allclose(t[:, 0, 0], b.field("T"), rtol=1e-11, atol=1e-300)


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED, L.POISSON_PROJECTED])
def test_multigrid_modes_close_to_oracle(emul, mode):
    """V-cycles on the deck case until the backward error and the last
    correction are both within 1e-10: every field within 1e-9 relative of the
    oracle's direct solve (the partitioned direct mode's level on this case),
    the interface within 1e-12, and where the cycle's vectors live (all in
    LDS, split, all in global memory) does not change a bit of the result."""
    case = L.deck_case()
    o = Oracle(case)
    assert o.step(8) == 0
    runs = []
    for cap in (-1, 3000, 0):
        b = L.Batch([case], lib=emul)
        b.set_poisson(mode, 1e-10, 100)
        b.set_mg_lds(cap)
        b.run(8)
        s = b.scalars()
        assert s["POISSON_SOLVES"] == 24
        assert 0 < s["POISSON_RESID"] <= 1e-10
        assert s["POISSON_CYCLES"] / s["POISSON_SOLVES"] <= (2 if mode == L.POISSON_MG64 else 5)
        zo = o.scalar_array()[3]
        assert abs(s["PLACE"] - zo) <= 1e-12 * zo
        rel = max_rel(b, 0, o, ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI"))
        assert all(v <= 1e-9 for v in rel.values()), rel
        runs.append([b.field(f) for f in ("W", "P", "T", "PHI")])
    for other in runs[1:]:
        assert all(np.array_equal(x, y) for x, y in zip(runs[0], other))


def _dense(a, b, c):
    return np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)


@pytest.mark.parametrize("mode", [L.POISSON_MG64, L.POISSON_MIXED])
@pytest.mark.parametrize("n", [1, 20, 33, 257, 1000])
def test_batched_poisson_multigrid_cpu(emul, mode, n):
    """lica_poisson by V-cycles against a dense solve (one level for n <= 32,
    odd level sizes above), on random diagonally dominant M-matrices of both
    sign patterns: the forward error within 1e-12 of the solution's maximum."""
    rng = np.random.default_rng(n)
    nsys = 3
    for sign in (1.0, -1.0):
        a = -sign * rng.uniform(0.5, 1.5, (nsys, n))
        c = -sign * rng.uniform(0.5, 1.5, (nsys, n))
        a[:, 0] = 0.0
        b = -(a + c) + sign * rng.uniform(0.01, 0.1, (nsys, n))
        g = rng.standard_normal((nsys, n))
        p = L.poisson(a, b, c, g, mode=mode, rtol=1e-10, max_cycles=100, lib=emul)
        for s in range(nsys):
            x = np.linalg.solve(_dense(a[s], b[s], c[s]), g[s])
            assert np.max(np.abs(p[s] - x)) <= 1e-12 * np.max(np.abs(x))


@pytest.mark.parametrize("mode,tol", [(L.POISSON_MG64, 1e-14), (L.POISSON_MIXED, 1e-11)])
def test_poisson_multigrid_nearly_singular_rows(emul, mode, tol):
    """The POISSON sign pattern with rows that sum to their rounding error
    (b = -(a + c) in floating point, the interior rows of 1d.f:3535-3541),
    a density jump of 1605 (1/rho of the vapour and the liquid) and the
    Dirichlet row on top: the solution within tol of an extended-precision
    elimination.  (A normwise backward error of 1e-10 alone would allow an
    error of order 1e-2 here, and TRDIAG3 in FP64 is off by about 2e-8 on
    these rows because its pivots lose the row sums.)"""
    a, b, c, g = nearly_singular_poisson()
    p = L.poisson(a, b, c, g, mode=mode, rtol=1e-10, max_cycles=100, lib=emul)[0]
    x = thomas_longdouble(a, b, c, g)
    assert np.max(np.abs(p - x)) <= tol * np.max(np.abs(x))


def test_poisson_multigrid_nonfinite_rows(emul):
    """A system with a non-finite right-hand side gives a non-finite solution
    (as the direct solve does) and is never reported converged with a finite
    residual."""
    n = 300
    a = np.full(n, 1.0)
    c = np.full(n, 1.0)
    a[0] = 0.0
    b = -(a + c) - 0.01
    g = np.ones(n)
    g[17] = np.nan
    p = L.poisson(a, b, c, g, mode=L.POISSON_MIXED, rtol=1e-10, max_cycles=10, lib=emul)
    assert not np.all(np.isfinite(p))


def test_projected_warm_start_cpu(emul):
    """The history projection (LICA_POISSON_PROJECTED) on the deck case: every
    solve meets the stopping test, the interface stays within 1e-12 relative
    of the oracle, and once the history holds the pressure field built up
    over the first steps the V-cycles per solve fall below the zero start's
    (four on this case)."""
    case = L.deck_case()
    o = Oracle(case)
    assert o.step(30) == 0
    cyc = {}
    for mode in (L.POISSON_MIXED, L.POISSON_PROJECTED):
        b = L.Batch([case], lib=emul)
        b.set_poisson(mode, 1e-10, 100)
        b.run(20)
        s0 = b.scalars()
        b.run(10)
        s = b.scalars()
        assert 0 < s["POISSON_RESID"] <= 1e-10
        zo = o.scalar_array()[3]
        assert abs(s["PLACE"] - zo) <= 1e-12 * zo
        cyc[mode] = (s["POISSON_CYCLES"] - s0["POISSON_CYCLES"]) / (s["POISSON_SOLVES"] - s0["POISSON_SOLVES"])
    assert cyc[L.POISSON_MIXED] >= 3.5
    assert cyc[L.POISSON_PROJECTED] <= 3, cyc


def test_reference_seams_bit_exact(emul):
    """MAIN's loop through the seams with the reference's argument lists
    (arrays shaped (0:2,0:2,0:N3) in Fortran order) gives the oracle's
    fields and scalars bit for bit."""
    from parity_util import check_reference_main
    from conftest import GOLDEN
    import os
    seams = L.ReferenceSeams(os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"), lib=emul)
    assert check_reference_main(seams, Oracle(L.deck_case()), 3, L.ref_array) == []


@pytest.mark.parametrize("mode,tol", [(L.POISSON_THOMAS, 0.0), (L.POISSON_MG64, 1e-12)])
def test_reference_poisson_seam(emul, mode, tol):
    """POISSON(P,DIVGSUM,DENF_XI,DENF_YI,DENF_ZI,DENF_Y,DENF_C) on the inputs
    of the oracle's last substep: the matrix assembled from DENF_ZI and the
    geometry, P(1,1,1:N) the oracle's TRDIAG3 solution (bit for bit in
    Thomas mode) and P(1,1,0), P(1,1,N3) its end copies (1d.f:3565-3572)."""
    from conftest import GOLDEN
    from oracle.binding import trdiag3
    import os
    o = Oracle(L.deck_case())
    assert o.step(3) == 0
    dzi, dg = o.last_poisson_inputs()
    a, b, c, g, _ = o.last_poisson()
    x = trdiag3(a, b, c, g)
    seams = L.ReferenceSeams(os.path.join(GOLDEN, "lica.in"), os.path.join(GOLDEN, "grid.in"), lib=emul)
    seams.set_poisson(mode, 1e-10, 100)
    n = seams.n
    P, DZI = L.ref_array(n), L.ref_array(n)
    DZI[:, 1, 1] = dzi
    seams.POISSON(P, dg, None, None, DZI, None, None)
    col = P[:, 1, 1]
    assert np.max(np.abs(col[1:n + 1] - x)) <= tol * np.max(np.abs(x))
    assert col[0] == col[1] and col[n + 1] == col[n]


def test_lane_bins_hold_equal_work(emul):
    """The binned lane split: consecutive runs of the largest-first order,
    each lane's work (n + 512 rows per case) within two cases' work of the
    others', every lane's largest case recorded; the solving kernels get a
    thread per chunk of the lane's largest case outside the Thomas mode; the
    interleaved split gives lane l the l-th largest case."""
    cases = L.batched_cases(200, seed=3)
    b = L.Batch(cases, init=False, lib=emul)
    b.set_lanes(4)
    b.set_lane_split(True)
    b.set_poisson(L.POISSON_MIXED, 1e-10, 100)
    ns = sorted((int(c.n) for c in cases), reverse=True)
    tab = b.tuning()
    assert len(tab) == 4
    pos, works = 0, []
    for t in tab:
        assert t["cases"] >= 1 and t["nmax"] == ns[pos]
        works.append(sum(n + 512 for n in ns[pos:pos + t["cases"]]))
        assert t["threads"]["momentum"] >= -(-t["nmax"] // 8)
        assert t["threads"]["cont_energy"] >= -(-t["nmax"] // 8)
        pos += t["cases"]
    assert pos == len(cases)
    assert max(works) - min(works) <= 2 * (ns[0] + 512)
    b.set_lane_split(False)
    assert [t["nmax"] for t in b.tuning()] == ns[:4]
    b.close()


@pytest.mark.parametrize("ireset,ipzero,idtold", [(False, False, True), (True, True, False)])
def test_fld_restart_matches_oracle(emul, tmp_path, ireset, ipzero, idtold):
    """Restart from a field file (IREAD=1, ITRACKING=1): PREFLD's reads,
    LVSINIT from the band read, GRID_COUPLING(1).  The CPU logic-check
    build and the oracle -- each with its own parser of the file -- give the
    same fields at the restart and five steps later, bit for bit; writing
    the restarted case again gives the file back but for the step counter
    (MAIN's loop starts again at M=1)."""
    from oracle.binding import read_fld
    case = L.deck_case()
    a = L.Batch([case], lib=emul)
    a.run(5)
    f1, f2 = tmp_path / "fld000005", tmp_path / "again"
    a.write_fld(f1, 0, 5)
    b = L.Batch([case], lib=emul)
    assert b.read_fld(f1, 0, ireset=ireset, ipzero=ipzero, idtold=idtold) == 0
    o = Oracle(case, init=False)
    assert o.init_fld(f1, ireset=ireset, ipzero=ipzero, idtold=idtold) == 0
    fields = ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI")
    for f in fields:
        assert np.array_equal(b.field(f), o.field(f)), f
    d = read_fld(f1)
    s = b.scalars()
    assert s["TIME"] == (0.0 if ireset else d["time"])
    assert s["DT"] == (d["dt"] if idtold else case.dt0)
    assert s["VOL_ORI"] == d["vol"]
    if not ipzero:
        b.write_fld(f2, 0, 5)
        x, y = open(f1).read().split("\n"), open(f2).read().split("\n")
        assert len(x) == len(y) and [i for i in range(len(x)) if x[i] != y[i]] == [1]
    b.run(5)
    assert o.step(5) == 0
    for f in fields:
        assert np.array_equal(b.field(f), o.field(f)), f
    assert np.array_equal(b.scalar_array()[:9], o.scalar_array()[:9])


def test_fld_reader_formats_and_refusals(emul, tmp_path):
    """Fortran's ES20.12 drops the E of a three-digit exponent
    (-1.234567890123-100); both parsers read it.  Files that are not a field
    file of the case's grid -- another N3, U not zero, x/y planes not
    uniform, cut short -- are refused with LICA_IO (status 22) and leave
    the batch usable."""
    from oracle.binding import _fortran_real, read_fld
    assert _fortran_real("-1.234567890123-100") == -1.234567890123e-100
    assert _fortran_real(" 2.500000000000E+00") == 2.5
    case = L.deck_case()
    a = L.Batch([case], lib=emul)
    a.run(3)
    good = tmp_path / "good"
    a.write_fld(good, 0, 3)
    lines = open(good).read().split("\n")
    n3 = case.n + 1
    # T(0:2,0:2,0) are the first nine values of the last list: give them a
    # three-digit exponent in Fortran's form
    nt = 9 * (n3 + 1)
    first = len(lines) - 1 - (nt + 4) // 5
    tiny = "%20s" % "1.000000000000-100"
    lines[first] = tiny * 5
    lines[first + 1] = tiny * 4 + lines[first + 1][80:]
    fortran = tmp_path / "fortran"
    open(fortran, "w").write("\n".join(lines))
    b = L.Batch([case], lib=emul)
    assert b.read_fld(fortran, 0, ireset=False, idtold=True) == 0
    assert b.field("T")[0] == 1e-100
    assert read_fld(fortran)["t"][0] == 1e-100
    bad = {}
    other = L.Batch([L.deck_case(200, 2)], lib=emul)
    other.write_fld(tmp_path / "other", 0, 0)
    bad["other grid"] = tmp_path / "other"
    lines = open(good).read().split("\n")
    u = list(lines)
    u[4] = "%20.12E" % 1.0 + u[4][20:]
    bad["U"] = tmp_path / "u"
    open(bad["U"], "w").write("\n".join(u))
    bad["cut"] = tmp_path / "cut"
    open(bad["cut"], "w").write("\n".join(lines[:len(lines) // 2]))
    nu = list(lines)
    nu[first] = nu[first][:20] + "%20.12E" % 0.5 + nu[first][40:]
    bad["T not uniform in x/y"] = tmp_path / "nonuniform"
    open(bad["T not uniform in x/y"], "w").write("\n".join(nu))
    for why, path in bad.items():
        with pytest.raises(L.LicaError) as e:
            b.read_fld(path, 0)
        assert e.value.status == 22, why
    b.run(1)


def test_oracle_loop_threads_do_not_change_results():
    """The oracle's reference-style mode (one case, the pseudo-time
    extrapolation's per-cell loops shared by the OpenMP threads, maxima
    combined by fmax) gives the serial oracle's fields bit for bit."""
    case = L.batched_cases(4096)[7]
    a, b = Oracle(case), Oracle(case)
    b.set_loop_threads(True)
    assert a.step(3) == b.step(3)
    for f in ("W", "P", "T", "PSIF", "PHI", "PSI_CN"):
        assert np.array_equal(a.field(f), b.field(f)), f
    assert np.array_equal(a.scalar_array(), b.scalar_array())
