"""Pins the oracle (the CPU restatement that the GPU path is compared with)
to what the mathematics fixes, independently of the oracle's own code:
textbook algorithms in a different algebraic form, exact closed forms,
brute force, invariants of the model.  CPU only."""
import math

import numpy as np
import pytest

import lica_boiling1d_amd.workloads as W
import oracle
from oracle import Oracle
from oracle.binding import trdiag3, grid_nodes

lib = oracle.load()


# ---- TRDIAG3 (1d.f:3783-3812): a textbook tridiagonal solve --------------
@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 333])
def test_trdiag3_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2.0, n)
    g = rng.standard_normal(n)
    x = trdiag3(a, b, c, g)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)   # a(1), c(n) ignored
    ref = np.linalg.solve(A, g)
    assert np.max(np.abs(x - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_trdiag3_toeplitz_closed_form():
    """-x(k-1) + 2x(k) - x(k+1) = 0 with x(0)=0, x(n+1)=1 gives x(k)=k/(n+1)."""
    n = 50
    a, b, c = -np.ones(n), 2 * np.ones(n), -np.ones(n)
    g = np.zeros(n)
    g[-1] = 1.0
    x = trdiag3(a, b, c, g)
    assert np.allclose(x, np.arange(1, n + 1) / (n + 1), rtol=0, atol=1e-14)


# ---- HJWENO (two_phase.f:2477-2500) vs the textbook WENO5 form -----------
def weno5_textbook(v, eps):
    """Osher & Fedkiw (2003) eqs. 3.28-3.35 from the five first differences."""
    v1, v2, v3, v4, v5 = v
    p1 = v1 / 3 - 7 * v2 / 6 + 11 * v3 / 6
    p2 = -v2 / 6 + 5 * v3 / 6 + v4 / 3
    p3 = v3 / 3 + 5 * v4 / 6 - v5 / 6
    s1 = 13 / 12 * (v1 - 2 * v2 + v3) ** 2 + 1 / 4 * (v1 - 4 * v2 + 3 * v3) ** 2
    s2 = 13 / 12 * (v2 - 2 * v3 + v4) ** 2 + 1 / 4 * (v2 - v4) ** 2
    s3 = 13 / 12 * (v3 - 2 * v4 + v5) ** 2 + 1 / 4 * (3 * v3 - 4 * v4 + v5) ** 2
    a1, a2, a3 = 0.1 / (s1 + eps) ** 2, 0.6 / (s2 + eps) ** 2, 0.3 / (s3 + eps) ** 2
    return (a1 * p1 + a2 * p2 + a3 * p3) / (a1 + a2 + a3)


def test_hjweno_equals_textbook_weno5():
    """Jiang & Peng's form (the reference's) with eps equals the textbook form
    with eps/12, since its smoothness indicators are 12x the textbook's."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        v = rng.standard_normal(5) * rng.choice([1e-3, 1.0, 1e3])
        a, b, c, d = v[1] - v[0], v[2] - v[1], v[3] - v[2], v[4] - v[3]
        ref = 1 / 12 * (-v[1] + 7 * v[2] + 7 * v[3] - v[4]) - lib.orc_hjweno(a, b, c, d)
        tb = weno5_textbook(v, 1e-6 / 12)
        assert abs(ref - tb) <= 1e-12 * max(1.0, np.max(np.abs(v)))


def test_hjweno_vanishes_on_quadratics():
    """Equal second differences (a quadratic) leave only the central part."""
    for k in (0.0, 1.0, -3.5, 1e4):
        assert lib.orc_hjweno(k, k, k, k) == 0.0


# ---- GODUNOV (two_phase.f:2448-2474) vs the Godunov Hamiltonian -----------
@pytest.mark.parametrize("s", [1.0, -1.0])
def test_godunov_matches_eikonal_hamiltonian(s):
    """Squared result equals max(max(D-,0)^2, min(D+,0)^2) for s>0 and
    max(min(D-,0)^2, max(D+,0)^2) for s<0 (Rouy & Tourin), except on the
    shock branch where the reference picks by the sign of s(D+ + D-)."""
    vals = [-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0]
    for p in vals:
        for m in vals:
            g = lib.orc_godunov(s, p, m)
            if s > 0:
                h = max(max(m, 0.0) ** 2, min(p, 0.0) ** 2)
            else:
                h = max(min(m, 0.0) ** 2, max(p, 0.0) ** 2)
            if s * p < 0 < s * m:      # converging characteristics: larger of the two
                h = max(p * p, m * m) if s * (p + m) != 0 else p * p
                assert g * g == (m * m if s * (p + m) > 0 else p * p)
                continue
            assert g * g == pytest.approx(h, abs=0), (s, p, m)


# ---- CAL_PSIF_MTP cell formula (two_phase.f:2869-2920) -------------------
def test_psif_cell_is_linear_cut_volume():
    """On a flat column the liquid fraction of a cell cut by a linear level
    set is clamp(1/2 + phi/D3, 0, 1); brute force by sub-sampling too."""
    rng = np.random.default_rng(1)
    for _ in range(500):
        d3 = rng.uniform(1e-4, 1.0)
        phi = rng.uniform(-1.0, 1.0) * d3
        got = lib.orc_psif_cell(phi, d3)
        exact = min(max(0.5 + phi / d3, 0.0), 1.0)
        assert abs(got - exact) <= 1e-14
        z = (np.arange(20000) + 0.5) / 20000 - 0.5       # cell coordinate
        assert abs(got - np.mean(phi + d3 * z > 0)) <= 1e-4
    assert lib.orc_psif_cell(0.0, 0.0) == 0.5                # degenerate branches
    assert lib.orc_psif_cell(1e-3, 0.0) == 1.0
    assert lib.orc_psif_cell(float("nan"), float("nan")) == -1.0   # the 'HAS BUG' branch


# ---- geometry: grid.f UNIFORM printed ES23.15 ----------------------------
def test_grid_nodes_are_es23_15_roundtrip():
    for n, zl in ((256, 1.0), (8192, 1.0), (37, 0.3)):
        z = grid_nodes(n, zl)
        ref = np.array([float(f"{k * (zl / n):.15E}") for k in range(n + 1)])
        assert np.array_equal(z, ref)


# ---- initial state and model invariants -----------------------------------
def test_initial_band_is_signed_distance_and_psi_is_box_average():
    c = W.deck_case()
    o = Oracle(c)
    assert o.init_status == 0
    n, r = int(c.n), int(c.r)
    band = o.field("BAND")
    alphi = o.field("ALPHI")
    planes, pos = [], 0
    for _ in range(11):
        cnt = int(band[pos]); planes += [int(k) for k in band[pos + 1:pos + 1 + cnt]]; pos += 1 + cnt
    dz = 1.0 / (r * n)
    zpf = (np.array(planes) - 0.5) * dz
    assert np.allclose(alphi, zpf - c.z0 + 1e-12, rtol=0, atol=1e-14)
    # PSIF is the liquid fraction of each refined cell of the initial plane
    psif = o.field("PSIF")
    zc = (np.arange(1, r * n + 1) - 0.5) * dz
    frac = np.clip(0.5 + (zc - c.z0 + 1e-12) / dz, 0.0, 1.0)
    frac = np.where(frac < 1e-5, 0.0, np.where(frac > 0.99999, 1.0, frac))
    assert np.allclose(psif, frac, rtol=0, atol=1e-12)
    # PSI_CN is the NS-cell average of PSIF (GRID_COUPLING(1)), snapped at 1e-3/0.999
    pcn = o.field("PSI_CN")[1:n + 1]
    avg = psif.reshape(n, r).mean(axis=1)
    avg = np.where(avg < 1e-3, 0.0, np.where(avg > 0.999, 1.0, avg))
    assert np.allclose(pcn, avg, rtol=0, atol=1e-14)


def test_no_superheat_no_motion():
    """Ja = 0: no interfacial mass flux, the film stays where it is and the
    level set stays the signed distance."""
    c = W.deck_case().copy(ja=0.0)
    o = Oracle(c)
    assert o.step(30) == 0
    s = o.scalar_array()
    assert abs(s[3] - c.z0) < 2e-6          # interface position
    band = o.field("BAND")
    alphi = o.field("ALPHI")
    planes, pos = [], 0
    for _ in range(11):
        cnt = int(band[pos]); planes += [int(k) for k in band[pos + 1:pos + 1 + cnt]]; pos += 1 + cnt
    zpf = (np.array(planes) - 0.5) / (int(c.r) * int(c.n))
    assert np.max(np.abs(alphi - (zpf - c.z0))) < 5e-6


def test_film_grows_monotonically():
    """With superheat the vapour film only grows (the sign of the interfacial
    mass flux, TF_GRAD chphase.f:1129-1130, and of the level-set source,
    two_phase.f:1469-1479).  The reference's source divides by the local
    phase density, so the growth comes in bursts when the band shifts and is
    slower than the Stefan similarity solution; the test asks for the sign
    and for growth over the run, not for the Stefan rate."""
    c = W.deck_case()
    o = Oracle(c)
    z = [c.z0]
    for _ in range(12):
        assert o.step(100) == 0
        z.append(o.scalar_array()[3])
    z = np.array(z)
    assert np.all(np.diff(z[1:]) >= 0.0)
    assert z[-1] > z[1] > c.z0


def test_wall_temperature_boundary_and_mass_bookkeeping():
    c = W.deck_case()
    o = Oracle(c)
    assert o.step(60) == 0
    T = o.field("T")
    n = int(c.n)
    assert T[0] == 0.0 and T[n] == 1.0 and T[n + 1] == 1.0     # LHSPS BCs (1d.f:4487-4494)
    s = o.scalar_array()
    assert s[6] > 0 and s[7] > s[6] * 0.99                       # vapour volume does not shrink


def test_rk3_coefficients_are_third_order():
    """The oracle's low-storage RK3 constants (RK3COEF 1d.f:1314-1319), read
    from the oracle: applied to y' = z y as the substeps use them
    (u_k = u_{k-1} + dt (gamma_k N_{k-1} + rho_k N_{k-2})), the amplification
    factor is exactly 1 + z + z^2/2 + z^3/6 (third order in three stages),
    and the implicit weights 2 alpha_k = gamma_k + rho_k sum to one."""
    g, r = Oracle(W.deck_case()).rk3_coefficients()
    P = np.polynomial.polynomial
    z = np.array([0.0, 1.0])
    u0 = np.array([1.0])
    n0 = P.polymul(z, u0)
    u1 = P.polyadd(u0, g[0] * n0)
    n1 = P.polymul(z, u1)
    u2 = P.polyadd(P.polyadd(u1, g[1] * n1), r[1] * n0)
    n2 = P.polymul(z, u2)
    u3 = P.polyadd(P.polyadd(u2, g[2] * n2), r[2] * n1)
    assert r[0] == 0.0
    assert np.allclose(u3[:4], [1.0, 1.0, 0.5, 1.0 / 6.0], rtol=0, atol=1e-15)
    assert math.isclose(float(np.sum(g + r)), 1.0, abs_tol=1e-15)


# ---- the box averages (INTER_PSI_FACE 1d.f:2552-2959 and relatives) ------
def test_box_average_equals_exact_integral():
    """The volume-weighted average of a piecewise-constant refined field over
    a coarse box equals its exact integral over the box divided by the box
    length: random boxes and the NS cells' own boxes, random fields."""
    c = W.deck_case()
    o = Oracle(c)
    n, r = o.n, o.r
    z = grid_nodes(n, c.zl)
    zf = grid_nodes(r * n, c.zl)
    rng = np.random.default_rng(1)
    f = rng.standard_normal(r * n)
    for t in range(1500):
        if t % 3 == 0:
            k = int(rng.integers(1, n + 1))
            z1, z2 = z[k - 1], z[k]
        else:
            z1, z2 = np.sort(rng.uniform(0.0, c.zl, 2))
        k1 = max(int(np.searchsorted(zf, z1, side="right")), 1)
        k2 = min(max(int(np.searchsorted(zf, z2, side="left")), k1), r * n)
        overlap = np.clip(np.minimum(z2, zf[1:]) - np.maximum(z1, zf[:-1]), 0.0, None)
        exact = float(np.sum(f * overlap) / (z2 - z1))
        got = o.box_avg(f, z1, z2, k1, k2)
        assert abs(got - exact) <= 1e-12 * max(abs(exact), 1.0), (t, z1, z2, got, exact)


# ---- DETERMINE_DT (1d.f:683-817) and CFL (1325-1373) ---------------------
def _dt_case():
    c = W.deck_case()
    return c, Oracle(c)


def test_determine_dt_growth_limit():
    """No velocity and no interface cell: the step may grow by 1.2 at most."""
    c, o = _dt_case()
    w = np.zeros(o.n + 2)
    psi = np.zeros(o.n + 2)
    psi[60:] = 1.0
    dt, cflm = o.determine_dt(w, psi, 2.5e-6)
    assert cflm == 0.0 and dt == 1.2 * 2.5e-6


def test_determine_dt_convective_limit():
    """Uniform velocity, no interface cell (the capillary limit is then 1000
    time units): the new step's convective CFL number, |w| dt / dz with w
    the cell-centre velocity, is CFLMAX, to the (dt/1000)/CFL^2 = 3e-6 that
    the capillary term of the CFL sum still adds."""
    c, o = _dt_case()
    wc = 3.0
    w = np.full(o.n + 2, wc)
    psi = np.zeros(o.n + 2)
    dz = 1.0 / o.n
    dt0 = 0.4 * dz / wc          # the current step's CFL is 0.4
    dt, cflm = o.determine_dt(w, psi, dt0)
    assert math.isclose(cflm, 0.4, rel_tol=1e-12)
    assert math.isclose(wc * dt / dz, c.cflmax, rel_tol=1e-5)
    assert wc * dt / dz < c.cflmax


def test_determine_dt_capillary_limit():
    """No velocity, one interface cell (0 < PSI_CN < 1): the capillary-gravity
    wave limit of Brackbill, Kothe & Zemach (J. Comput. Phys. 100, 1992,
    eq. 61) with the gravity-wave term, dt_s = dx / (2 c) with
    c^2 = pi sigma / ((rho_l + rho_v) dx) + g (rho_l - rho_v) dx / ((rho_l + rho_v) pi),
    sigma = 1/We and g = 1/Fr^2 in the deck's units, dx the cell's smaller
    extent; the step update dt' = CFLMAX sqrt(dt dt_s), the geometric mean
    the CFL-sum rule gives without convection.  A surface tension four times
    larger halves dt_s."""
    c, o = _dt_case()
    w = np.zeros(o.n + 2)
    psi = np.zeros(o.n + 2)
    psi[40] = 0.3
    dx = min(1.0, 1.0 / o.n)          # lateral extent YL = 1, dz = 1/n
    rl, rv = c.denr, 1.0

    def dt_s(sigma):
        c2 = math.pi * sigma / ((rl + rv) * dx) + (rl - rv) / c.fr ** 2 * dx / ((rl + rv) * math.pi)
        return dx / (2.0 * math.sqrt(c2))

    dt0 = 1e-7
    dt, cflm = o.determine_dt(w, psi, dt0)
    assert cflm == 0.0
    assert math.isclose(dt, c.cflmax * math.sqrt(dt0 * dt_s(1.0 / c.surf_j)), rel_tol=1e-12)
    o4 = Oracle(c.copy(surf_j=c.surf_j / 4.0))
    dt4, _ = o4.determine_dt(w, psi, dt0)
    assert math.isclose(dt4 / dt, math.sqrt(dt_s(4.0 / c.surf_j) / dt_s(1.0 / c.surf_j)), rel_tol=1e-12)
    assert math.isclose(dt4 / dt, 1.0 / math.sqrt(2.0), rel_tol=1e-9)


# ---- GLOBAL_MASS_CORRECTION (two_phase.f:6196-6318) ----------------------
def _band_planes(o):
    band = o.field("BAND")
    out, pos = [], 0
    for nl in range(1, 12):
        cnt = int(band[pos])
        out += [(nl, int(k)) for k in band[pos + 1:pos + 1 + cnt]]
        pos += 1 + cnt
    return out


def _crossing(o, c):
    """zero crossing of the band's level set (linear between refined centres)"""
    zpf = (np.arange(1, o.r * o.n + 1) - 0.5) / (o.r * o.n) * c.zl
    phi = dict((k, v) for (nl, k), v in zip(_band_planes(o), o.field("ALPHI")))
    ks = sorted(phi)
    for a, b in zip(ks[:-1], ks[1:]):
        if b == a + 1 and phi[a] < 0.0 <= phi[b]:
            return zpf[a - 1] + (zpf[b - 1] - zpf[a - 1]) * (-phi[a]) / (phi[b] - phi[a]), phi
    raise AssertionError("no crossing")


@pytest.mark.parametrize("dv", [1e-5, 3e-4, -2e-3])
def test_mass_correction_postcondition(dv):
    """On the initial signed distance (whose vapour volume per unit area is
    the interface height), asking for a vapour volume dv larger: the level
    set of layers 1..NN1 shifts by one constant, the outer layers stay, the
    interface moves toward the target, and when fewer than 15 passes were
    made the volume ends within 1e-5 of the target."""
    c = W.deck_case()
    o = Oracle(c)
    v0 = o.scalar_array()[6]
    z0, phi0 = _crossing(o, c)
    assert abs(z0 - v0) < 1e-9
    layer = dict((k, nl) for nl, k in _band_planes(o))
    st, passes = o.mass_correction(v0 + dv)
    assert st == 0 and 1 <= passes <= 15
    z1, phi1 = _crossing(o, c)
    shift = np.array([phi1[k] - phi0[k] for k in phi0 if layer[k] <= 8])
    outer = np.array([phi1[k] - phi0[k] for k in phi0 if layer[k] > 8])
    assert np.ptp(shift) <= 1e-15 and np.all(outer == 0.0)
    assert math.isclose(z1 - z0, -shift[0], rel_tol=1e-9)     # phi = z - z_i: the crossing moves by -shift
    assert (z1 - z0) * dv > 0.0 and abs(v0 + dv - z1) < abs(dv)
    if passes < 15:
        assert abs(v0 + dv - z1) < 1e-5


# ---- VEL_INTP / VELASSIGN / VEL_R3 (two_phase.f:5539-6191) ----------------
def test_velocity_interpolation_reproduces_cubics():
    """The 4-point Lagrange interpolation of the cell-centre velocity to the
    refined faces is exact for a cubic: W chosen so that every cell-centre
    average is a cubic, the refined velocity of the inner band layers equals
    the cubic at the refined faces."""
    c = W.deck_case()
    o = Oracle(c)
    n, r = o.n, o.r
    z = grid_nodes(n, c.zl)
    zf = grid_nodes(r * n, c.zl)
    centre = np.concatenate([[z[0] - 0.5 * (z[1] - z[0])], 0.5 * (z[:-1] + z[1:])])   # cells 0..n

    def cub(x):
        return 0.3 - 1.7 * x + 2.2 * x ** 2 + 4.1 * x ** 3

    w = np.zeros(n + 2)
    w[0] = cub(z[0])
    for k in range(n + 1):
        w[k + 1] = 2.0 * cub(centre[k]) - w[k]
    wf = o.vel_intp(w)
    inner = [k for nl, k in _band_planes(o) if nl <= 8]
    assert len(inner) >= 16
    for k in inner:
        assert abs(wf[k - 1] - cub(zf[k - 1])) <= 1e-13 * np.max(np.abs(w)), k


# ---- TEMP_GRAD: TEMP_EXTP, TF_GRAD, VOL_TOT_MFDOT (chphase.f:55-1493) ------
def _piecewise_linear_t(o, c, sv, sl):
    z = grid_nodes(o.n, c.zl)
    zc = np.concatenate([[z[0] - 0.5 * (z[1] - z[0])], 0.5 * (z[:-1] + z[1:]), [z[-1] + 0.5 * (z[-1] - z[-2])]])
    t = np.where(zc < c.z0, c.tsat + sv * (zc - c.z0), c.tsat + sl * (zc - c.z0))
    return zc, t


def test_temperature_extension_fixed_point_and_stefan_flux():
    """A temperature linear on each side of the interface (vapour slope sv,
    liquid slope sl, T = T_sat at the interface): each phase's extension
    across the interface continues its own line (the linear profile is the
    fixed point of the constant-normal-derivative extension), and the mass
    flux is the Stefan condition mdot = Ja/(Re Pr) (k_v sv - k_l sl) with the
    deck's k_v = 1 and k_l = TCR -- to 2e-4 on the two layers next to the
    interface and 1e-2 out to layer 6, as far as the extension's RAT < 1e-4
    convergence test takes it."""
    c = W.deck_case()
    o = Oracle(c)
    sv, sl = 16.0, -3.0
    zc, t = _piecewise_linear_t(o, c, sv, sl)
    st, t0, t1, mfl, sf2, place, vol = o.temp_grad(t, 1e-6)
    assert st == 0
    near = np.abs(zc - c.z0) < 4.0 / o.n
    dzc = 1.0 / o.n
    assert np.max(np.abs(t0 - (c.tsat + sv * (zc - c.z0)))[near]) <= 0.02 * sv * dzc
    assert np.max(np.abs(t1 - (c.tsat + sl * (zc - c.z0)))[near]) <= 0.02 * abs(sl) * dzc
    mdot = c.ja / (c.re * c.prm) * (1.0 * sv - c.tcr * sl)
    planes = _band_planes(o)
    for (nl, k), m in zip(planes, mfl):
        if nl <= 2:
            assert abs(m - mdot) <= 2e-4 * mdot, (nl, k, m, mdot)
        elif nl <= 6:
            assert abs(m - mdot) <= 1e-2 * mdot, (nl, k, m, mdot)


def test_provisional_volume_change_follows_the_mass_flux():
    """VOL_TOT_MFDOT's provisional level-set advance is, after its RK blend
    (first stage 2 DTCONST, then 3/4-1/4 and 1/3-2/3 averages), one forward
    Euler step phi -= DTCONST mdot / rho with rho the density of the plane's
    own phase (two_phase.f:1469-1479 form); the zero crossing of that
    advance predicts the provisional interface and the vapour volume change
    to within the displacement's own 15 % (the reinitialisation that follows
    moves the crossing by about a tenth of it), while a wrong density side,
    sign or blend would miss by a factor of two or more.  The volume change
    equals the sum of SFGAS2 (the NS source's volume per unit area) exactly."""
    c = W.deck_case()
    o = Oracle(c)
    sv, sl = 16.0, -3.0
    zc, t = _piecewise_linear_t(o, c, sv, sl)
    v0 = o.scalar_array()[6]
    z0, phi0 = _crossing(o, c)
    dtc = 1e-6
    st, t0, t1, mfl, sf2, place, vol = o.temp_grad(t, dtc)
    assert st == 0
    mdot = c.ja / (c.re * c.prm) * (1.0 * sv - c.tcr * sl)
    zpf = (np.arange(1, o.r * o.n + 1) - 0.5) / (o.r * o.n) * c.zl
    moved = {k: p - dtc * mdot / (c.denr if p > 0.0 else 1.0) for k, p in phi0.items()}
    ks = sorted(moved)
    zp = None
    for a, b in zip(ks[:-1], ks[1:]):
        if b == a + 1 and moved[a] < 0.0 <= moved[b]:
            zp = zpf[a - 1] + (zpf[b - 1] - zpf[a - 1]) * (-moved[a]) / (moved[b] - moved[a])
    dz_pred = zp - z0
    assert dz_pred > 0.0
    assert abs((place - z0) - dz_pred) <= 0.15 * dz_pred
    assert abs((vol - v0) - dz_pred) <= 0.15 * dz_pred
    assert math.isclose(float(np.sum(sf2)) * c.zl / (o.r * o.n), vol - v0, rel_tol=1e-12)


# ---- DIVGS, POISSON, UCALC and CONVRGE1 (1d.f:3180-3627, 3306-3362) -------
@pytest.mark.parametrize("which,steps", [("deck", 30), ("batch1", 10)])
def test_projection_enforces_the_phase_change_divergence(which, steps):
    """After a step the projected velocity satisfies the continuity equation
    with phase change, dw/dz = S (rho_l - rho_v)/(rho_l rho_v) on every cell
    2..N-1 (S the cell's box average of SFGAS), to 2e-9 of the source's
    size (the POISSON solve's rounding; 3e-11 on the deck, 5e-10 on the
    nearly motionless column of batch case 1, whose w is 1e-7); cell 1 keeps the residue of the reference's unscaled first row
    (1d.f:3295-3299), and CONVRGE1's DVMAX is the largest |dw/dz - S (rho_l -
    rho_v)/rho_l| over the cells."""
    c = W.deck_case() if which == "deck" else W.batched_cases(4096)[1]
    o = Oracle(c)
    assert o.step(steps) == 0
    n, r = o.n, o.r
    z = grid_nodes(n, c.zl)
    zf = grid_nodes(r * n, c.zl)
    wv, sf = o.field("W"), o.field("SFGAS")
    sbar = np.array([np.sum(sf * np.clip(np.minimum(z[k], zf[1:]) - np.maximum(z[k - 1], zf[:-1]), 0.0, None))
                     / (z[k] - z[k - 1]) for k in range(1, n + 1)])
    div = (wv[2:n + 1] - wv[1:n]) / (z[1:n] - z[:n - 1])      # cells 1..n-1
    src = sbar[:n - 1] * (c.denr - 1.0) / (c.denr * 1.0)
    assert np.count_nonzero(src) >= 1
    assert np.max(np.abs(div - src)[1:]) <= 2e-9 * np.max(np.abs(src))
    dvmax = np.max(np.abs(div - sbar[:n - 1] * (c.denr - 1.0) / c.denr))
    assert math.isclose(o.scalar_array()[4], dvmax, rel_tol=1e-12)
