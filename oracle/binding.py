"""ctypes binding of oracle/lica_oracle.h (TEST INFRASTRUCTURE ONLY)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liblica_oracle.so")

ORC_FIELDS = {"W": 0, "P": 1, "T": 2, "PSI_CN": 3, "PSI_ZN": 4, "PSIF": 5, "PHI": 6,
              "SFGAS": 7, "WO": 8, "BAND": 9, "ALPHI": 10, "MFLVS": 11}
NSCAL = 24
c_i64, c_dbl = ctypes.c_int64, ctypes.c_double
P_dbl = ctypes.POINTER(ctypes.c_double)


class OrcCase(ctypes.Structure):
    _fields_ = [("n", c_i64), ("r", c_i64), ("zl", c_dbl),
                ("re", c_dbl), ("visr", c_dbl), ("denr", c_dbl),
                ("surf_j", c_dbl), ("fr", c_dbl),
                ("prm", c_dbl), ("scr", c_dbl), ("tcr", c_dbl),
                ("tsat", c_dbl), ("ja", c_dbl),
                ("dt0", c_dbl), ("cflmax", c_dbl),
                ("idtopt", c_i64),
                ("z0", c_dbl), ("tslope", c_dbl)]


def build(force=False, extra=()):
    """gcc -O3 -fopenmp -ffp-contract=off: every operation rounds on its own,
    in the serial order the reference's loops run; OpenMP only spreads
    independent cases over threads (orc_run_cases)."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "lica_oracle.c")
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= max(
            os.path.getmtime(src), os.path.getmtime(os.path.join(HERE, "lica_oracle.h"))):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off", "-std=c11", "-Wall", *extra, src,
           "-o", tmp, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        V = ctypes.c_void_p
        lib.orc_new.restype = V
        lib.orc_new.argtypes = [ctypes.POINTER(OrcCase)]
        lib.orc_free.argtypes = [V]
        lib.orc_init.argtypes = [V]
        lib.orc_init.restype = ctypes.c_int
        lib.orc_step.argtypes = [V, c_i64]
        lib.orc_step.restype = ctypes.c_int
        lib.orc_get.argtypes = [V, ctypes.c_int, P_dbl, c_i64]
        lib.orc_get.restype = c_i64
        lib.orc_scalars.argtypes = [V, P_dbl]
        lib.orc_n.argtypes = [V]
        lib.orc_n.restype = c_i64
        lib.orc_last_poisson.argtypes = [V, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]
        lib.orc_last_poisson.restype = c_i64
        lib.orc_trdiag3.argtypes = [c_i64, P_dbl, P_dbl, P_dbl, P_dbl]
        lib.orc_hjweno.argtypes = [c_dbl] * 4
        lib.orc_hjweno.restype = c_dbl
        lib.orc_godunov.argtypes = [c_dbl] * 3
        lib.orc_godunov.restype = c_dbl
        lib.orc_psif_cell.argtypes = [c_dbl, c_dbl]
        lib.orc_psif_cell.restype = c_dbl
        lib.orc_grid_nodes.argtypes = [c_i64, c_dbl, P_dbl]
        P_i64 = ctypes.POINTER(c_i64)
        lib.orc_run_cases.argtypes = [ctypes.POINTER(OrcCase), c_i64, c_i64, P_i64, P_i64, P_dbl, P_dbl]
        lib.orc_run_cases.restype = ctypes.c_int
        lib.orc_rk3_coefficients.argtypes = [V, P_dbl, P_dbl]
        lib.orc_test_box_avg.argtypes = [V, P_dbl, c_dbl, c_dbl, c_i64, c_i64]
        lib.orc_test_box_avg.restype = c_dbl
        lib.orc_test_determine_dt.argtypes = [V, P_dbl, P_dbl, c_dbl, P_dbl, P_dbl]
        lib.orc_test_mass_correction.argtypes = [V, c_dbl, P_i64]
        lib.orc_test_mass_correction.restype = ctypes.c_int
        lib.orc_test_vel_intp.argtypes = [V, P_dbl, P_dbl]
        lib.orc_test_temp_grad.argtypes = [V, P_dbl, c_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]
        lib.orc_test_temp_grad.restype = ctypes.c_int
        lib.orc_last_poisson_inputs.argtypes = [V, P_dbl, P_dbl]
        lib.orc_last_poisson_inputs.restype = c_i64
        lib.orc_set_loop_threads.argtypes = [V, ctypes.c_int]
        lib.orc_run_single.argtypes = [ctypes.POINTER(OrcCase), c_i64, c_i64, P_i64, P_i64, P_dbl, P_dbl]
        lib.orc_run_single.restype = ctypes.c_int
        lib.orc_init_fld.argtypes = [V, P_dbl, P_dbl, P_dbl, c_dbl, c_dbl, c_dbl, P_i64, P_i64, P_dbl]
        lib.orc_init_fld.restype = ctypes.c_int
        lib.orc_trdiag_lines.argtypes = [P_dbl] * 5 + [c_i64] * 6 + [ctypes.c_int]
        lib.orc_trdiag_lines.restype = None
        lib.orc_trdiag_periodic_lines.argtypes = [P_dbl] * 4 + [c_i64] * 6
        lib.orc_trdiag_periodic_lines.restype = None
        lib.orc_ptdiag_lines.argtypes = [P_dbl] * 7 + [c_i64] * 6 + [P_dbl] * 3
        lib.orc_ptdiag_lines.restype = None
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(P_dbl)


ORC_M_MAX, ORC_MAXP = 11, 16


def _fortran_real(t):
    t = t.strip()
    if not t:
        return 0.0
    u = t.upper().replace("D", "E")
    if "E" not in u:
        # a three-digit exponent written without its E (1.0-100)
        j = max(u.rfind("+"), u.rfind("-"))
        if j > 0:
            u = u[:j] + "E" + u[j:]
    return float(u)


class _Records:
    """Fortran formatted READs of a text file: every READ starts a new line,
    a list longer than its format continues on the next one."""

    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().split("\n")
        self.i = 0

    def line(self):
        s = self.lines[self.i]
        self.i += 1
        return s

    def read(self, widths, kinds, count=None):
        """One READ: `count` values (default len(widths)) in fields of the
        given widths and kinds ('i' or 'r'), the format reused per line."""
        count = len(widths) if count is None else count
        out = []
        while True:
            s = self.line()
            at = 0
            for w, kd in zip(widths, kinds):
                if len(out) == count:
                    break
                fld = s[at:at + w]
                at += w
                out.append(int(fld) if kd == "i" and fld.strip() else (0 if kd == "i" else _fortran_real(fld)))
            if len(out) >= count:
                return out


def read_fld(path):
    """The lists PREFLD reads from a field file (1d.f:972-1097), for the
    1-D column: N1..N3, IHIST, M, TIME, DT, IPSS, W(k), P(k), T(k),
    VOL_TOT_ORI and the band as ordered planes per layer (each the r*r
    list entries of one z-plane)."""
    r = _Records(path)
    n1, n2, n3 = r.read([8] * 3, "iii")
    ihist, m, time, dt = r.read([8, 8, 20, 20], "iirr")
    ipss = r.read([8, 8, 8, 8, 20], "iiiir")[0]
    r.read([20], "r")
    u = r.read([20] * 5, "r" * 5, n1 * (n2 + 1) * (n3 + 1))
    v = r.read([20] * 5, "r" * 5, (n1 + 1) * n2 * (n3 + 1))
    w = r.read([20] * 5, "r" * 5, (n1 + 1) * (n2 + 1) * n3)
    p = r.read([20] * 5, "r" * 5, (n1 - 1) * (n2 - 1) * (n3 - 1))
    n1f, n2f, n3f = r.read([8] * 3, "iii")
    r.read([20] * 6, "r" * 6)
    nlvs, nmax = r.read([8, 8], "ii")
    vol = r.read([20] * 8, "r" * 8, nlvs)
    numa = r.read([8] * 8, "i" * 8, nmax * nlvs)
    rr = n1f - 1
    band = []
    for nl in range(nmax):
        planes = []
        for q in range(numa[nl] // (rr * rr)):
            cells = [r.read([8, 8, 8, 20], "iiir") for _ in range(rr * rr)]
            assert all(c[2] == cells[0][2] and c[3] == cells[0][3] for c in cells)
            planes.append((cells[0][2], cells[0][3]))
        band.append(planes)
    r.read([8] * 4, "i" * 4, nlvs * nlvs)        # CONTACT
    r.read([8] * 4, "i" * 4, nlvs * nlvs)        # CONTACT_OLD
    r.read([20] * 5, "r" * 5, nlvs * nlvs)       # CONTACT_START
    r.read([20] * 5, "r" * 5, nlvs * nlvs)       # CONTACT_TIME
    r.read([8] * 4, "i" * 4, nlvs)               # LVSON
    r.read([20] * 5, "r" * 5, nlvs * nlvs)       # T_CON
    t = None
    if ipss == 1:
        r.read([20] * 3, "rrr")
        t = r.read([20] * 5, "r" * 5, (n1 + 1) * (n2 + 1) * (n3 + 1))
    per = (n1 + 1) * (n2 + 1)
    col = lambda a, nk: np.array([a[k * per] for k in range(nk)])
    return {"n3": n3, "n3f": n3f, "ihist": ihist, "m": m, "time": time, "dt": dt, "ipss": ipss,
            "u0": not any(u) and not any(v),
            "w": col(w, n3), "p": np.array(p), "t": col(t, n3 + 1) if t is not None else np.zeros(n3 + 1),
            "vol": vol[0], "band": band}


class Oracle:
    """One case of the serial FP64 restatement."""

    def __init__(self, case, init=True):
        self.lib = load()
        oc = OrcCase()
        for f, _ in OrcCase._fields_:
            setattr(oc, f, getattr(case, f))
        self.case = oc
        self.h = self.lib.orc_new(ctypes.byref(oc))
        self.n = int(case.n)
        self.r = int(case.r)
        self.init_status = self.lib.orc_init(self.h) if init else None

    def close(self):
        if self.h:
            self.lib.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1):
        return self.lib.orc_step(self.h, int(nsteps))

    def set_loop_threads(self, on=True):
        """Share the per-cell loops of the pseudo-time extrapolation among
        the OpenMP threads (the reference's style within one case)."""
        self.lib.orc_set_loop_threads(self.h, 1 if on else 0)

    def init_fld(self, path, ireset=True, ipzero=False, idtold=False):
        """In place of init: MAIN with IREAD=1, ITRACKING=1 from a field
        file (PREFLD, LVSINIT from the band read, GRID_COUPLING(1)), with
        the deck's IRESET / IPZERO / IDTOLD."""
        d = read_fld(path)
        n3 = d["n3"]
        w = np.zeros(n3 + 1)
        w[1:] = d["w"]
        p = np.zeros(n3 + 1)
        if not ipzero:
            p[1:n3] = d["p"]
        t = np.array(d["t"], dtype=np.float64)
        bn = np.zeros(ORC_M_MAX + 1, dtype=np.int64)
        bk = np.zeros((ORC_M_MAX + 1, ORC_MAXP), dtype=np.int64)
        al = np.zeros((ORC_M_MAX + 1, ORC_MAXP))
        for nl, planes in enumerate(d["band"], start=1):
            bn[nl] = len(planes)
            for q, (k, a) in enumerate(planes):
                bk[nl, q] = k
                al[nl, q] = a
        P_i64 = ctypes.POINTER(c_i64)
        return self.lib.orc_init_fld(self.h, _p(w), _p(p), _p(t), 0.0 if ireset else d["time"],
                                     d["dt"] if idtold else float(self.case.dt0), d["vol"],
                                     bn.ctypes.data_as(P_i64), bk.ctypes.data_as(P_i64), _p(al))

    def field(self, name):
        f = ORC_FIELDS[name] if isinstance(name, str) else int(name)
        cap = max(self.n + 2, self.r * self.n, 12 * 11 * 17) + 16
        out = np.zeros(cap)
        m = self.lib.orc_get(self.h, f, _p(out), cap)
        return out[:m].copy()

    def scalar_array(self):
        out = np.zeros(NSCAL)
        self.lib.orc_scalars(self.h, _p(out))
        return out

    # unit-level hooks (pinning tests)
    def rk3_coefficients(self):
        g, r = np.zeros(3), np.zeros(3)
        self.lib.orc_rk3_coefficients(self.h, _p(g), _p(r))
        return g, r

    def box_avg(self, f, z1, z2, k1, k2):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == self.n * self.r
        return self.lib.orc_test_box_avg(self.h, _p(f), float(z1), float(z2), int(k1), int(k2))

    def determine_dt(self, w, psi_cn, dt):
        w, psi_cn = (np.ascontiguousarray(x, dtype=np.float64) for x in (w, psi_cn))
        assert w.size == self.n + 2 and psi_cn.size == self.n + 2
        dto, cflm = ctypes.c_double(), ctypes.c_double()
        self.lib.orc_test_determine_dt(self.h, _p(w), _p(psi_cn), float(dt), ctypes.byref(dto), ctypes.byref(cflm))
        return dto.value, cflm.value

    def mass_correction(self, vol_target):
        passes = ctypes.c_int64()
        st = self.lib.orc_test_mass_correction(self.h, float(vol_target), ctypes.byref(passes))
        return st, passes.value

    def vel_intp(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.n + 2
        wf = np.zeros(self.n * self.r)
        self.lib.orc_test_vel_intp(self.h, _p(w), _p(wf))
        return wf

    def temp_grad(self, t, dtconst):
        t = np.ascontiguousarray(t, dtype=np.float64)
        assert t.size == self.n + 2
        t0, t1 = np.zeros(self.n + 2), np.zeros(self.n + 2)
        mfl = np.zeros(11 * 16)
        sf2 = np.zeros(self.n * self.r)
        place, vol = ctypes.c_double(), ctypes.c_double()
        st = self.lib.orc_test_temp_grad(self.h, _p(t), float(dtconst), _p(t0), _p(t1), _p(mfl), _p(sf2),
                                         ctypes.byref(place), ctypes.byref(vol))
        return st, t0, t1, mfl, sf2, place.value, vol.value

    def last_poisson_inputs(self):
        """DENF_ZI (0..n+1) and DIVGSUM (1..n) of the last substep's POISSON"""
        dzi, dg = np.zeros(self.n + 2), np.zeros(self.n)
        self.lib.orc_last_poisson_inputs(self.h, _p(dzi), _p(dg))
        return dzi, dg

    def last_poisson(self):
        n = self.n
        a, b, c, g, p = (np.zeros(n) for _ in range(5))
        self.lib.orc_last_poisson(self.h, _p(a), _p(b), _p(c), _p(g), _p(p))
        return a, b, c, g, p


def trdiag3(a, b, c, g):
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64).copy() for x in (a, b, c))
    g = np.ascontiguousarray(g, dtype=np.float64).copy()
    load().orc_trdiag3(len(g), _p(a), _p(b), _p(c), _p(g))
    return g


def _fa(x):
    """a column-major float64 copy (the reference's array layout)"""
    return np.array(x, dtype=np.float64, order="F", copy=True)


def trdiag_lines(a, b, c, r, l1, l2, ll1, ll2, overwrite_r=False):
    """TRDIAG1/2/3 (1d.f:3721-3812) on m1 x m2 arrays (1-based ranges):
    returns (uu, r) -- r changes only with overwrite_r (TRDIAG3)."""
    a, b, c, r = (_fa(x) for x in (a, b, c, r))
    uu = np.zeros_like(r, order="F")
    m1, m2 = r.shape
    load().orc_trdiag_lines(_p(a), _p(b), _p(c), _p(r), _p(uu), m1, m2, l1, l2, ll1, ll2, int(overwrite_r))
    return uu, r


def trdiag_periodic_lines(a, b, c, f, j1, j2, l1, l2):
    """TRDIAG1P/2P/3P (1d.f:3815-3985): returns F after the solve."""
    a, b, c, f = (_fa(x) for x in (a, b, c, f))
    m1, m2 = f.shape
    load().orc_trdiag_periodic_lines(_p(a), _p(b), _p(c), _p(f), m1, m2, j1, j2, l1, l2)
    return f


def ptdiag_lines(a, b, c, d, e, f, i1, i_n, j1, jn):
    """PTDIAG1A (1d.f:3988-4034): returns (uu, o, q, r)."""
    a, b, c, d, e, f = (_fa(x) for x in (a, b, c, d, e, f))
    m1, m2 = f.shape
    uu, o, q, r = (np.zeros_like(f, order="F") for _ in range(4))
    load().orc_ptdiag_lines(_p(a), _p(b), _p(c), _p(d), _p(e), _p(f), _p(uu), m1, m2, i1, i_n, j1, jn, _p(o), _p(q),
                            _p(r))
    return uu, o, q, r


def grid_nodes(n, zl):
    z = np.zeros(n + 1)
    load().orc_grid_nodes(n, zl, _p(z))
    return z


def run_cases(cases, nsteps):
    """Set up and advance every case nsteps steps (or until it stops), one
    case per OpenMP thread.  Returns (status, steps entered, set-up seconds,
    stepping seconds, threads)."""
    lib = load()
    arr = (OrcCase * len(cases))()
    for i, c in enumerate(cases):
        for f, _ in OrcCase._fields_:
            setattr(arr[i], f, getattr(c, f))
    st = np.zeros(len(cases), dtype=np.int64)
    steps = np.zeros(len(cases), dtype=np.int64)
    t_set, t_step = ctypes.c_double(), ctypes.c_double()
    P_i64 = ctypes.POINTER(c_i64)
    thr = lib.orc_run_cases(arr, len(cases), int(nsteps), st.ctypes.data_as(P_i64), steps.ctypes.data_as(P_i64),
                            ctypes.byref(t_set), ctypes.byref(t_step))
    return st, steps, t_set.value, t_step.value, int(thr)


def run_single(cases, nsteps, threads=0):
    """The reference's style: the cases one after the other, each with its
    per-cell extrapolation loops shared by `threads` OpenMP threads (0: the
    default count).  Returns (status, steps entered, set-up seconds,
    stepping seconds, threads)."""
    lib = load()
    st = np.zeros(len(cases), dtype=np.int64)
    steps = np.zeros(len(cases), dtype=np.int64)
    t_set = t_step = 0.0
    thr = 1
    P_i64 = ctypes.POINTER(c_i64)
    for i, c in enumerate(cases):
        oc = OrcCase()
        for f, _ in OrcCase._fields_:
            setattr(oc, f, getattr(c, f))
        s1, s2 = ctypes.c_double(), ctypes.c_double()
        thr = lib.orc_run_single(ctypes.byref(oc), int(nsteps), int(threads), st[i:].ctypes.data_as(P_i64),
                                 steps[i:].ctypes.data_as(P_i64), ctypes.byref(s1), ctypes.byref(s2))
        t_set += s1.value
        t_step += s2.value
    return st, steps, t_set, t_step, int(thr)
