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
