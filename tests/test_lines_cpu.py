"""The line solvers of the 3-D code (1d.f:3721-4034): the oracle against dense
solves, and the library's entry points (CPU logic-check build) against the
oracle bit for bit.  Arrays are (m1, m2), the systems along the second index,
1-based ranges as in the reference."""
import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd import _abi as abi
from oracle import binding as O

import emul_build


@pytest.fixture(scope="module")
def emul():
    return abi.load(emul_build.build())


def _rng(seed):
    return np.random.default_rng(seed)


def _tri_inputs(m1, m2, seed, dominant=True):
    g = _rng(seed)
    a, c, r = (g.uniform(-1.0, 1.0, (m1, m2)) for _ in range(3))
    b = (3.0 if dominant else 0.5) + g.uniform(0.0, 1.0, (m1, m2))
    b *= np.where(g.uniform(size=(m1, m2)) < 0.5, -1.0, 1.0)
    return a, b, c, r


def _penta_inputs(m1, m2, seed):
    g = _rng(seed)
    b, c, d, e, f = (g.uniform(-1.0, 1.0, (m1, m2)) for _ in range(5))
    a = 5.0 + g.uniform(0.0, 1.0, (m1, m2))
    return a, b, c, d, e, f


SHAPES = [(1, 2, 1, 2, 1, 1), (7, 9, 2, 8, 1, 7), (33, 17, 1, 17, 5, 30), (64, 40, 3, 38, 2, 63)]


@pytest.mark.parametrize("m1,m2,l1,l2,ll1,ll2", SHAPES)
def test_oracle_trdiag_equals_dense_solves(m1, m2, l1, l2, ll1, ll2):
    """Each line's rows l1..l2 solve the tridiagonal system A u(j-1) +
    B u(j) + C u(j+1) = R (A(l1), C(l2) outside it) to rounding; lines
    outside ll1..ll2 and rows outside l1..l2 keep UU's zeros."""
    a, b, c, r = _tri_inputs(m1, m2, 1)
    uu, r_out = O.trdiag_lines(a, b, c, r, l1, l2, ll1, ll2)
    assert np.array_equal(r_out, r)
    n = l2 - l1 + 1
    for i in range(ll1 - 1, ll2):
        M = np.diag(b[i, l1 - 1:l2]) + np.diag(a[i, l1:l2], -1) + np.diag(c[i, l1 - 1:l2 - 1], 1)
        x = np.linalg.solve(M, r[i, l1 - 1:l2])
        assert np.max(np.abs(uu[i, l1 - 1:l2] - x)) <= 1e-13 * max(1.0, np.max(np.abs(x))), i
    mask = np.zeros((m1, m2), bool)
    mask[ll1 - 1:ll2, l1 - 1:l2] = True
    assert np.all(uu[~mask] == 0.0)
    assert n >= 1


def test_oracle_trdiag3_leaves_the_solution_in_r():
    """TRDIAG3 (1d.f:3804-3808) stores UU into R on the backward sweep,
    rows l1..l2-1; R(l2) keeps the right-hand side."""
    a, b, c, r = _tri_inputs(9, 12, 2)
    uu, r3 = O.trdiag_lines(a, b, c, r, 2, 11, 3, 8, overwrite_r=True)
    uu1, _ = O.trdiag_lines(a, b, c, r, 2, 11, 3, 8)
    assert np.array_equal(uu, uu1)
    assert np.array_equal(r3[2:8, 1:10], uu[2:8, 1:10])
    assert np.array_equal(r3[2:8, 10], r[2:8, 10])
    keep = np.ones_like(r, bool)
    keep[2:8, 1:10] = False
    assert np.array_equal(r3[keep], r[keep])


@pytest.mark.parametrize("m1,m2,j1,j2,l1,l2", [(1, 2, 1, 2, 1, 1), (5, 8, 1, 8, 1, 5), (20, 31, 4, 29, 3, 18)])
def test_oracle_periodic_equals_dense_cyclic_solves(m1, m2, j1, j2, l1, l2):
    """TRDIAG1P: the cyclic system whose first row couples to the last
    unknown through A(j1) and whose last row couples to the first through
    C(j2), solved to rounding; rows and lines outside the ranges unchanged."""
    a, b, c, f = _tri_inputs(m1, m2, 3)
    out = O.trdiag_periodic_lines(a, b, c, f, j1, j2, l1, l2)
    n = j2 - j1 + 1
    for k in range(l1 - 1, l2):
        M = np.zeros((n, n))
        for t in range(n):
            j = j1 - 1 + t
            M[t, t] = b[k, j]
            M[t, (t - 1) % n] += a[k, j]
            M[t, (t + 1) % n] += c[k, j]
        x = np.linalg.solve(M, f[k, j1 - 1:j2])
        assert np.max(np.abs(out[k, j1 - 1:j2] - x)) <= 1e-13 * max(1.0, np.max(np.abs(x))), k
    mask = np.zeros((m1, m2), bool)
    mask[l1 - 1:l2, j1 - 1:j2] = True
    assert np.array_equal(out[~mask], f[~mask])


@pytest.mark.parametrize("m1,m2,i1,i_n,j1,jn", [(1, 2, 1, 1, 1, 2), (6, 3, 1, 6, 1, 3), (17, 25, 2, 15, 3, 24)])
def test_oracle_pentadiagonal_equals_dense_solves(m1, m2, i1, i_n, j1, jn):
    """PTDIAG1A: E u(j-2) + D u(j-1) + A u(j) + B u(j+1) + C u(j+2) = F on
    rows j1..jn, solved to rounding (terms reaching outside the rows drop)."""
    a, b, c, d, e, f = _penta_inputs(m1, m2, 4)
    uu, o, q, r = O.ptdiag_lines(a, b, c, d, e, f, i1, i_n, j1, jn)
    n = jn - j1 + 1
    for i in range(i1 - 1, i_n):
        M = np.zeros((n, n))
        for t in range(n):
            j = j1 - 1 + t
            for off, coef in ((-2, e), (-1, d), (0, a), (1, b), (2, c)):
                if 0 <= t + off < n:
                    M[t, t + off] = coef[i, j]
        x = np.linalg.solve(M, f[i, j1 - 1:jn])
        assert np.max(np.abs(uu[i, j1 - 1:jn] - x)) <= 1e-12 * max(1.0, np.max(np.abs(x))), i
    # the back substitution the coefficients describe: u = O u+ + Q u++ + R
    for i in range(i1 - 1, i_n):
        for j in range(j1 - 1, jn - 2):
            v = o[i, j] * uu[i, j + 1] + q[i, j] * uu[i, j + 2] + r[i, j]
            assert v == uu[i, j]


@pytest.mark.parametrize("m1,m2,l1,l2,ll1,ll2", SHAPES)
def test_library_trdiag_bit_exact(emul, m1, m2, l1, l2, ll1, ll2):
    """lica_trdiag1/2/3 (CPU logic-check build) give the oracle's UU and R
    bit for bit; UU's elements outside the ranges keep the caller's values."""
    a, b, c, r = _tri_inputs(m1, m2, 5)
    uu0 = _rng(6).uniform(-1, 1, (m1, m2))
    for kind in (1, 2, 3):
        uu, r_out = L.trdiag(kind, a, b, c, r, l1, l2, ll1, ll2, uu=uu0, lib=emul)
        ref_uu, ref_r = O.trdiag_lines(a, b, c, r, l1, l2, ll1, ll2, overwrite_r=(kind == 3))
        mask = np.zeros((m1, m2), bool)
        mask[ll1 - 1:ll2, l1 - 1:l2] = True
        assert np.array_equal(uu[mask], ref_uu[mask]), kind
        assert np.array_equal(uu[~mask], uu0[~mask]), kind
        assert np.array_equal(r_out, ref_r), kind


def test_library_periodic_and_pentadiagonal_bit_exact(emul):
    """lica_trdiag1p/2p/3p and lica_ptdiag1a (CPU logic-check build) give
    the oracle's results bit for bit."""
    a, b, c, f = _tri_inputs(23, 19, 7)
    ref = O.trdiag_periodic_lines(a, b, c, f, 2, 18, 3, 21)
    for kind in (1, 2, 3):
        assert np.array_equal(L.trdiag_periodic(kind, a, b, c, f, 2, 18, 3, 21, lib=emul), ref), kind
    a, b, c, d, e, f = _penta_inputs(21, 16, 8)
    ref = O.ptdiag_lines(a, b, c, d, e, f, 2, 20, 2, 15)
    out = L.ptdiag1a(a, b, c, d, e, f, 2, 20, 2, 15, lib=emul)
    for x, y in zip(out, ref):
        assert np.array_equal(x, y)


def test_library_line_solvers_refuse_bad_ranges(emul):
    """Ranges outside the arrays, empty ranges and the one-row periodic or
    pentadiagonal systems (whose reference loops read outside their rows)
    are refused with LICA_BAD_ARG."""
    a, b, c, r = _tri_inputs(4, 5, 9)
    for args in ((0, 5, 1, 4), (1, 6, 1, 4), (1, 5, 1, 5), (3, 2, 1, 4), (1, 5, 3, 2)):
        with pytest.raises(L.LicaError):
            L.trdiag(1, a, b, c, r, *args, lib=emul)
    with pytest.raises(L.LicaError):
        L.trdiag_periodic(1, a, b, c, r, 2, 2, 1, 4, lib=emul)
    p = _penta_inputs(4, 5, 10)
    with pytest.raises(L.LicaError):
        L.ptdiag1a(*p, 1, 4, 3, 3, lib=emul)
