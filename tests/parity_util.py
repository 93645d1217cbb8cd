"""Comparison helpers of the parity tests (library vs oracle)."""
import numpy as np

from oracle import Oracle

# the fields and scalars both sides expose, in the order they are checked
FIELDS = ("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI", "SFGAS", "WO", "BAND", "ALPHI", "MFLVS")
# TIME, DT, STEP, PLACE, DVMAX, CFLM, VOL_ORI, VOL_IPHS, QWALL, STATUS, counters
NSCAL_SHARED = 17


def exact_mismatches(batch, inst, orc):
    """Names of fields/scalars whose values differ in any bit."""
    bad = []
    for f in FIELDS:
        g, c = batch.field(f, inst), orc.field(f)
        if g.shape != c.shape or not np.array_equal(g, c):
            bad.append(f)
    s, so = batch.scalar_array(inst), orc.scalar_array()
    for i in range(NSCAL_SHARED):
        if s[i] != so[i]:
            bad.append(f"scalar{i}")
    return bad


def max_rel(batch, inst, orc, fields=("W", "P", "T", "PSI_CN", "PSI_ZN", "PSIF", "PHI")):
    """max |a-b| / max(|b|_inf, tiny) per field."""
    out = {}
    for f in fields:
        g, c = batch.field(f, inst), orc.field(f)
        if g.shape != c.shape:
            out[f] = np.inf
            continue
        scale = max(np.max(np.abs(c)), 1e-300)
        out[f] = float(np.max(np.abs(g - c)) / scale) if g.size else 0.0
    return out


def ulp_sensitivity(case, steps, fields):
    """max relative change of each field when the oracle's own input z0
    moves by one ulp: the part of a field that rounding decides."""
    o1, o2 = Oracle(case), Oracle(case.copy(z0=np.nextafter(case.z0, 1.0)))
    o1.step(steps)
    o2.step(steps)
    out = {}
    for f in fields:
        x, y = o1.field(f), o2.field(f)
        out[f] = float(np.max(np.abs(x - y)) / max(np.max(np.abs(x)), 1e-300))
    o1.close()
    o2.close()
    return out


def reference_main(seams, steps, ref_array):
    """MAIN's time loop (1d.f:196-295) through the seams with the reference's
    argument lists, on arrays shaped like the reference's: DETERMINE_DT, the
    zeroing of RK3ZO/ANPSO (1d.f:206-218), then per substep RHSNLHS,
    CONVRGE1, TRANSPORT and GRID_COUPLING(0), and TIME = TIME + DT."""
    n = seams.n
    st = seams.state()
    w, p, zn, cn = st["W"], st["P"], st["PSI_ZN"], st["PSI_CN"]
    vol, dt = st["VOL_TOT_ORI"], st["DT"]
    u, v, xn, yn, rx, ry, rz, an = (ref_array(n) for _ in range(8))
    dvmax = cflm = 0.0
    for _ in range(steps):
        cflm, dt = seams.DETERMINE_DT(u, v, w, cn, dt)
        rz[:] = 0.0
        an[:] = 0.0
        for msub in (1, 2, 3):
            seams.RHSNLHS(msub, 2, u, v, w, p, xn, yn, zn, cn, rx, ry, rz, an)
            dvmax = seams.CONVRGE1(u, v, w)
            vol = seams.TRANSPORT(u, v, w, cn, vol)
            seams.GRID_COUPLING(0, xn, yn, zn, cn)
        seams.end_step()
    untouched = all(not np.any(x) for x in (u, v, xn, yn, rx, ry))
    return {"W": w, "P": p, "PSI_ZN": zn, "PSI_CN": cn, "VOL_TOT_ORI": vol, "DVMAX": dvmax, "CFLM": cflm,
            "DT": dt, "untouched": untouched}


def check_reference_main(seams, orc, steps, ref_array):
    """The seam-driven run against the oracle, bit for bit: the caller's
    columns I = J = 1, the other columns left zero, the context's own state."""
    out = reference_main(seams, steps, ref_array)
    assert orc.step(steps) == 0
    bad = []
    for f in ("W", "P", "PSI_ZN", "PSI_CN"):
        arr = out[f]
        if not np.array_equal(arr[:, 1, 1], orc.field(f)):
            bad.append(f)
        rest = arr.copy()
        rest[:, 1, 1] = 0.0
        if np.any(rest):
            bad.append(f + " outside the column")
    so = orc.scalar_array()
    for name, i in (("DVMAX", 4), ("VOL_TOT_ORI", 6), ("DT", 1), ("CFLM", 5)):
        if out[name] != so[i]:
            bad.append(name)
    if not out["untouched"]:
        bad.append("x/y arrays written")
    return bad + exact_mismatches(seams, 0, orc)


def thomas_longdouble(a, b, c, g):
    """Thomas elimination in x87 extended precision (64-bit significand)."""
    a, b, c, g = (np.asarray(v, dtype=np.longdouble) for v in (a, b, c, g))
    n = len(g)
    cp, gp = np.zeros(n, np.longdouble), np.zeros(n, np.longdouble)
    cp[0], gp[0] = c[0] / b[0], g[0] / b[0]
    for k in range(1, n):
        m = b[k] - a[k] * cp[k - 1]
        cp[k] = c[k] / m if k < n - 1 else 0.0
        gp[k] = (g[k] - a[k] * gp[k - 1]) / m
    x = np.zeros(n, np.longdouble)
    x[-1] = gp[-1]
    for k in range(n - 2, -1, -1):
        x[k] = gp[k] - cp[k] * x[k + 1]
    return x.astype(np.float64)


def nearly_singular_poisson(seed=5, n=3000, interface=900, ratio=1605.36):
    """POISSON rows (1d.f:3535-3541) as the batch's hardest systems have
    them: b = -(a + c) in floating point (interior rows sum to their
    rounding error), 1/rho jumping by the liquid/vapour density ratio at the
    interface row, Neumann at the wall, and the last row's c (the implicit
    p = 0 above the top face) kept in its b only."""
    rng = np.random.default_rng(seed)
    rho = np.where(np.arange(n + 1) < interface, 1.0, ratio)
    coef = (1.0 / rho) * rng.uniform(0.99, 1.01, n + 1)
    a = coef[:-1].copy()
    c = coef[1:].copy()
    a[0] = 0.0
    b = -(a + c)
    g = rng.standard_normal(n) * 1e3
    return a, b, c, g

