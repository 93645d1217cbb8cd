"""Host mirror of the reference's driver calls over the C ABI.

``Context`` follows MAIN's call sequence for one case (1d.f:70-295):
``determine_dt`` (DETERMINE_DT 1d.f:683), ``rhsnlhs`` (RHSNLHS 1d.f:1444,
which contains the POISSON projection 1d.f:3403-3488), ``convrge1``
(CONVRGE1 1d.f:3306), ``transport`` (TRANSPORT two_phase.f:701 followed by
GRID_COUPLING(0) two_phase.f:4863, as at 1d.f:242-244).  ``Batch`` runs the
same loop for many independent cases in one launch per kernel.
Non-zero statuses raise ``LicaError`` (the reference STOPs).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import CaseDesc, LicaError, FIELDS, SCALARS, NSCAL


def _check(rc, where):
    if rc != 0:
        raise LicaError(rc, where)


def _dptr(a):
    return a.ctypes.data_as(_abi.P_dbl)


def parse_deck(lica_in, grid_in, lib=None):
    """lica.in + grid.in -> CaseDesc (INPUT_SIMULATION 1d.f:323-492, grid.f:29-53)."""
    lib = lib if lib is not None else _abi.load()
    d = CaseDesc()
    _check(lib.lica_parse_deck(str(lica_in).encode(), str(grid_in).encode(), ctypes.byref(d)), "lica_parse_deck")
    return d


class _Handle:
    def __init__(self, lib=None):
        self._lib = lib if lib is not None else _abi.load()
        self._h = ctypes.c_void_p()

    def close(self):
        if self._h:
            self._lib.lica_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_poisson(self, mode, rtol=1.0e-10, max_cycles=100):
        _check(self._lib.lica_batch_set_poisson(self._h, int(mode), float(rtol), int(max_cycles)), "set_poisson")

    # getters shared by Batch and Context
    def field(self, name, inst=0):
        f = FIELDS[name] if isinstance(name, str) else int(name)
        n = int(self.cases[inst].n)
        r = int(self.cases[inst].r)
        cap = max(n + 2, r * n, 12 * 11 * 17) + 16
        out = np.zeros(cap)
        cnt = ctypes.c_int64()
        _check(self._lib.lica_batch_get_field(self._h, f, inst, _dptr(out), cap, ctypes.byref(cnt)), "get_field")
        return out[:cnt.value].copy()

    def scalars(self, inst=0):
        out = np.zeros(NSCAL)
        _check(self._lib.lica_batch_get_scalars(self._h, inst, _dptr(out)), "get_scalars")
        return {k: float(out[i]) for i, k in enumerate(SCALARS)}

    def scalar_array(self, inst=0):
        out = np.zeros(NSCAL)
        _check(self._lib.lica_batch_get_scalars(self._h, inst, _dptr(out)), "get_scalars")
        return out


class Batch(_Handle):
    """Many independent boiling cases advanced together on the GPU."""

    def __init__(self, cases, init=True, lib=None, _handle=None):
        super().__init__(lib)
        self.cases = list(cases)
        if _handle is not None:
            self._h = _handle
            return
        arr = (CaseDesc * len(self.cases))(*self.cases)
        _check(self._lib.lica_batch_create(len(self.cases), arr, ctypes.byref(self._h)), "lica_batch_create")
        if init:
            self.init()

    def __len__(self):
        return len(self.cases)

    def save(self, path):
        """Binary checkpoint of the whole device state (bit-exact restart)."""
        _check(self._lib.lica_batch_save(self._h, str(path).encode()), "lica_batch_save")

    @classmethod
    def load(cls, path, lib=None):
        lib = lib if lib is not None else _abi.load()
        h = ctypes.c_void_p()
        _check(lib.lica_batch_load(str(path).encode(), ctypes.byref(h)), "lica_batch_load")
        n = int(lib.lica_batch_size(h))
        # the case list is read back from the file header by the library; keep
        # a Python copy for the getters' sizes
        import struct
        with open(path, "rb") as f:
            f.read(8)
            hdr = struct.unpack("<6q", f.read(48))
            raw = f.read(ctypes.sizeof(CaseDesc) * n)
        cases = [CaseDesc.from_buffer_copy(raw, i * ctypes.sizeof(CaseDesc)) for i in range(n)]
        return cls(cases, lib=lib, _handle=h)

    def write_fld(self, path, inst=0, ihist=0):
        """WRITEFIELD text file of one case (1d.f:1121-1232)."""
        _check(self._lib.lica_batch_write_fld(self._h, int(inst), str(path).encode(), int(ihist)), "write_fld")

    def read_fld(self, path, inst=0, ireset=True, ipzero=False, idtold=False):
        """Restart case `inst` from a field file as MAIN does with IREAD=1,
        ITRACKING=1 (PREFLD 1d.f:972-1097, LVSINIT two_phase.f:146-189,
        GRID_COUPLING(1)); ireset/ipzero/idtold are the deck's IRESET,
        IPZERO, IDTOLD (lica.in line 2; the shipped deck has 1, 0, 0).
        Returns the case's status (0, or a reference STOP)."""
        flags = (1 if ireset else 0) | (2 if ipzero else 0) | (4 if idtold else 0)
        rc = self._lib.lica_batch_read_fld(self._h, int(inst), str(path).encode(), flags)
        if rc != 0 and not _abi.is_case_status(rc):
            _check(rc, "read_fld")
        return int(rc)

    def init(self):
        _check(self._lib.lica_batch_init(self._h), "lica_batch_init")

    def run(self, nsteps):
        _check(self._lib.lica_batch_run(self._h, int(nsteps)), "lica_batch_run")

    def set_mg_lds(self, nbytes):
        """Cap on the multigrid kernel's LDS (bytes; -1 restores the default)."""
        _check(self._lib.lica_batch_set_mg_lds(self._h, int(nbytes)), "set_mg_lds")

    def set_block(self, threads):
        _check(self._lib.lica_batch_set_block(self._h, int(threads)), "set_block")

    def use_graph(self, on=True):
        _check(self._lib.lica_batch_use_graph(self._h, 1 if on else 0), "use_graph")

    def set_net(self, net):
        """Network of the learned warm start (LICA_POISSON_LEARNED)."""
        self._net = net
        _check(self._lib.lica_batch_set_net(self._h, net.handle), "set_net")

    def set_capture(self, on=True, substep=3):
        """Keep the POISSON system of substep `substep` (1..3) of every step."""
        _check(self._lib.lica_batch_set_capture(self._h, int(substep) if on else 0), "set_capture")

    def set_lanes(self, lanes=2):
        """Run the cases as `lanes` sets (1..8) on as many streams at once."""
        _check(self._lib.lica_batch_set_lanes(self._h, int(lanes)), "set_lanes")

    def set_lane_split(self, binned=True):
        """Lanes as contiguous size bins of the largest-first order holding
        equal work, each launched with shapes fitted to its own largest case
        (binned), or as every L-th case of that order (the default)."""
        _check(self._lib.lica_batch_set_lane_split(self._h, 1 if binned else 0), "set_lane_split")

    def autotune(self, steps=2):
        """Coordinate descent on the step time: for every kernel class whose
        shape is free and every lane, 64..1024 threads per workgroup, then
        multigrid LDS sizes, each trial `steps` steps of all lanes at once;
        a candidate within 0.5 % of the best or faster is confirmed by trials
        of 4 x `steps` steps in turn with the kept shape, twice, and taken
        when more than 0.5 % faster there.  The batch returns to
        the state it had; no result depends on the choice.  Returns
        tuning()."""
        _check(self._lib.lica_batch_autotune(self._h, int(steps)), "autotune")
        return self.tuning()

    def tuning(self):
        """Launch shape per lane: cases, largest n, multigrid LDS bytes and
        threads per workgroup of each kernel class."""
        cnt = ctypes.c_int64(0)
        _check(self._lib.lica_batch_get_tuning(self._h, None, 0, ctypes.byref(cnt)), "get_tuning")
        out = np.zeros(cnt.value, dtype=np.int64)
        _check(self._lib.lica_batch_get_tuning(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), out.size,
                                               ctypes.byref(cnt)), "get_tuning")
        per = 3 + len(KERNEL_CLASSES)
        res = []
        for l in range(out.size // per):
            o = out[l * per:(l + 1) * per]
            res.append({"cases": int(o[0]), "nmax": int(o[1]), "mg_lds": int(o[2]),
                        "threads": {k: int(o[3 + i]) for i, k in enumerate(KERNEL_CLASSES) if k in TUNED_CLASSES}})
        return res

    def warm_info(self):
        """Per case [first row of the network window, backward error of the
        warm-start base, s_r, s_b, alpha, start-test backward error, 1 if the
        solve took the guess as it was] of the last learned-mode substep."""
        out = np.zeros((len(self.cases), 7))
        _check(self._lib.lica_batch_warm_info(self._h, _dptr(out), out.size), "warm_info")
        return out

    def set_timing(self, on=True):
        _check(self._lib.lica_batch_set_timing(self._h, 1 if on else 0), "set_timing")

    def sync(self):
        _check(self._lib.lica_batch_sync(self._h), "sync")

    def status(self):
        bad = ctypes.c_int64(-1)
        rc = self._lib.lica_batch_status(self._h, ctypes.byref(bad))
        return rc, bad.value

    def progress(self):
        """(status, completed steps) of every case."""
        st = np.zeros(len(self.cases), dtype=np.int64)
        steps = np.zeros(len(self.cases), dtype=np.int64)
        _check(self._lib.lica_batch_get_progress(self._h, st.ctypes.data_as(_abi.P_i64),
                                                 steps.ctypes.data_as(_abi.P_i64)), "get_progress")
        return st, steps

    def run_all(self, nsteps):
        """Advance every healthy case by nsteps; a case that stops (the
        reference STOPs there) keeps its status and is left behind.
        Returns the number of cases with a non-zero status."""
        rc = self._lib.lica_batch_run(self._h, int(nsteps))
        if rc != 0 and _abi.is_case_status(rc):
            # the status of the first stopped case (a reference STOP or a
            # Poisson solve at its cycle cap); the others are in progress()
            st, _ = self.progress()
            return int(np.count_nonzero(st))
        _check(rc, "lica_batch_run")
        return 0

    def timings(self):
        """GPU ms per kernel class of the last run (timing mode) and the wall ms."""
        out = np.zeros(16)
        m = self._lib.lica_batch_last_timings(self._h, _dptr(out), 16)
        # the work between the POISSON matrix and its solve (warm-start base,
        # capture, network) is timed in the slot after poisson_mg
        names = ["init", "dt", "cont_energy", "temp_grad", "momentum", "poisson", "ucalc", "transport", "poisson_mg",
                 "warm_start"]
        res = {names[i]: float(out[i]) for i in range(min(m - 1, len(names)))}
        res["wall"] = float(out[m - 1])
        return res


# kernel classes in the library's order, and those whose launch shape the
# autotuner sets
KERNEL_CLASSES = ["init", "dt", "cont_energy", "temp_grad", "momentum", "poisson_direct", "ucalc", "transport",
                  "poisson_mg", "capture_pre", "capture_post", "warm_base", "convrge1"]
TUNED_CLASSES = ("dt", "cont_energy", "momentum", "poisson_direct", "ucalc", "poisson_mg")


class Context(_Handle):
    """One case driven through the reference's subroutine sequence."""

    def __init__(self, lica_in, grid_in, lib=None):
        super().__init__(lib)
        self.cases = [parse_deck(lica_in, grid_in, lib=self._lib)]
        _check(self._lib.lica_init(ctypes.byref(self._h), str(lica_in).encode(), str(grid_in).encode()), "lica_init")
        self.n = int(self.cases[0].n)

    def step(self, nsteps=1):
        _check(self._lib.lica_step(self._h, int(nsteps)), "lica_step")

    def determine_dt(self):
        cflm, dt = ctypes.c_double(), ctypes.c_double()
        _check(self._lib.lica_determine_dt(self._h, ctypes.byref(cflm), ctypes.byref(dt)), "lica_determine_dt")
        return cflm.value, dt.value

    def rhsnlhs(self, msub):
        bufs = [np.zeros(self.n + 2) for _ in range(4)]
        _check(self._lib.lica_rhsnlhs(self._h, int(msub), *[_dptr(b) for b in bufs]), "lica_rhsnlhs")
        return tuple(bufs)

    def convrge1(self):
        d = ctypes.c_double()
        _check(self._lib.lica_convrge1(self._h, ctypes.byref(d)), "lica_convrge1")
        return d.value

    def transport(self):
        _check(self._lib.lica_transport(self._h), "lica_transport")

    def grid_coupling(self, ini_psi=0):
        zn, cn = np.zeros(self.n + 2), np.zeros(self.n + 2)
        _check(self._lib.lica_grid_coupling(self._h, int(ini_psi), _dptr(zn), _dptr(cn)), "lica_grid_coupling")
        return zn, cn

    def end_step(self):
        _check(self._lib.lica_end_step(self._h), "lica_end_step")


def ref_array(n):
    """A zeroed array shaped like the reference's REAL*8 (0:2,0:2,0:N3) with
    N3 = n+1, in Fortran memory order: numpy shape (N3+1, 3, 3) indexed
    [K, J, I]; the column the 1-D path uses is [:, 1, 1]."""
    return np.zeros((n + 2, 3, 3))


class ReferenceSeams(Context):
    """One case driven through the seams with the reference's own argument
    lists (DETERMINE_DT 1d.f:683, RHSNLHS 1d.f:1444, CONVRGE1 1d.f:3306,
    TRANSPORT two_phase.f:701, GRID_COUPLING two_phase.f:4863, POISSON
    1d.f:3492, TEMP_GRAD chphase.f:55) on arrays from ref_array: each call
    reads and writes the caller's arrays as the subroutine does."""

    def _a(self, x):
        if x is None:
            return None
        assert x.shape == (self.n + 2, 3, 3) and x.dtype == np.float64 and x.flags.c_contiguous
        return _dptr(x)

    def state(self):
        """W, P, T, PSI_ZN, PSI_CN arrays and VOL_TOT_ORI, TIME, DT of the context"""
        w, p, t, zn, cn = (ref_array(self.n) for _ in range(5))
        vol, tim, dt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self._lib.lica_f_state(self._h, *[self._a(x) for x in (w, p, t, zn, cn)], ctypes.byref(vol),
                                      ctypes.byref(tim), ctypes.byref(dt)), "lica_f_state")
        return {"W": w, "P": p, "T": t, "PSI_ZN": zn, "PSI_CN": cn, "VOL_TOT_ORI": vol.value, "TIME": tim.value,
                "DT": dt.value}

    def DETERMINE_DT(self, u, v, w, psi_cn, dt):
        cflm, d = ctypes.c_double(), ctypes.c_double(dt)
        _check(self._lib.lica_f_determine_dt(self._h, ctypes.byref(cflm), self._a(u), self._a(v), self._a(w),
                                             self._a(psi_cn), ctypes.byref(d)), "lica_f_determine_dt")
        return cflm.value, d.value

    def RHSNLHS(self, msub, itracking, u, v, w, p, psi_xn, psi_yn, psi_zn, psi_cn, rk3xo, rk3yo, rk3zo, anpso,
                qvol_ori=0.0, vol1=0.0):
        ms, it = ctypes.c_int64(msub), ctypes.c_int64(itracking)
        q, v1 = ctypes.c_double(qvol_ori), ctypes.c_double(vol1)
        arrs = [self._a(x) for x in (u, v, w, p, psi_xn, psi_yn, psi_zn, psi_cn, rk3xo, rk3yo, rk3zo, anpso)]
        _check(self._lib.lica_f_rhsnlhs(self._h, ctypes.byref(ms), ctypes.byref(it), *arrs,
                                        ctypes.byref(q), ctypes.byref(v1)), "lica_f_rhsnlhs")

    def CONVRGE1(self, u, v, w):
        d = ctypes.c_double()
        _check(self._lib.lica_f_convrge1(self._h, ctypes.byref(d), self._a(u), self._a(v), self._a(w)),
               "lica_f_convrge1")
        return d.value

    def TRANSPORT(self, u, v, w, psi_cn, vol_tot_ori):
        vol = ctypes.c_double(vol_tot_ori)
        _check(self._lib.lica_f_transport(self._h, self._a(u), self._a(v), self._a(w), self._a(psi_cn),
                                          ctypes.byref(vol)), "lica_f_transport")
        return vol.value

    def GRID_COUPLING(self, ini_psi, psi_xn, psi_yn, psi_zn, psi_cn):
        ip = ctypes.c_int64(ini_psi)
        _check(self._lib.lica_f_grid_coupling(self._h, ctypes.byref(ip), self._a(psi_xn), self._a(psi_yn),
                                              self._a(psi_zn), self._a(psi_cn)), "lica_f_grid_coupling")

    def POISSON(self, p, divgsum, denf_xi, denf_yi, denf_zi, denf_y, denf_c):
        divgsum = np.ascontiguousarray(divgsum, dtype=np.float64)
        assert divgsum.size == self.n
        _check(self._lib.lica_f_poisson(self._h, self._a(p), _dptr(divgsum), self._a(denf_xi), self._a(denf_yi),
                                        self._a(denf_zi), self._a(denf_y), self._a(denf_c)), "lica_f_poisson")

    def TEMP_GRAD(self, psi_cn):
        _check(self._lib.lica_f_temp_grad(self._h, self._a(psi_cn)), "lica_f_temp_grad")


def _fcol(x):
    """a column-major float64 copy: element (i, j) at (i-1) + (j-1)*m1, the
    reference's array layout"""
    return np.array(x, dtype=np.float64, order="F", copy=True)


def trdiag(kind, a, b, c, r, l1, l2, ll1, ll2, uu=None, lib=None):
    """TRDIAG1/2/3 (1d.f:3721-3812) on (m1, m2) arrays, 1-based ranges: the
    tridiagonal systems along the second index for lines ll1..ll2.  Returns
    (uu, r); r changes only for kind 3 (TRDIAG3 leaves UU in R(:, l1..l2-1))."""
    lib = lib if lib is not None else _abi.load()
    a, b, c, r = (_fcol(x) for x in (a, b, c, r))
    uu = np.zeros_like(r, order="F") if uu is None else _fcol(uu)
    m1, m2 = r.shape
    fn = {1: lib.lica_trdiag1, 2: lib.lica_trdiag2, 3: lib.lica_trdiag3}[int(kind)]
    _check(fn(_dptr(a), _dptr(b), _dptr(c), _dptr(r), _dptr(uu), m1, m2, l1, l2, ll1, ll2), "lica_trdiag")
    return uu, r


def trdiag_periodic(kind, a, b, c, f, j1, j2, l1, l2, lib=None):
    """TRDIAG1P/2P/3P (1d.f:3815-3985): periodic systems along the second
    index (rows j1..j2) for lines l1..l2; returns F after the solve."""
    lib = lib if lib is not None else _abi.load()
    a, b, c, f = (_fcol(x) for x in (a, b, c, f))
    m1, m2 = f.shape
    fn = {1: lib.lica_trdiag1p, 2: lib.lica_trdiag2p, 3: lib.lica_trdiag3p}[int(kind)]
    _check(fn(_dptr(a), _dptr(b), _dptr(c), _dptr(f), m1, m2, j1, j2, l1, l2), "lica_trdiag_p")
    return f


def ptdiag1a(a, b, c, d, e, f, i1, i_n, j1, jn, uu=None, o=None, q=None, r=None, lib=None):
    """PTDIAG1A (1d.f:3988-4034): pentadiagonal systems along the second
    index (rows j1..jn) for lines i1..i_n; returns (uu, o, q, r)."""
    lib = lib if lib is not None else _abi.load()
    a, b, c, d, e, f = (_fcol(x) for x in (a, b, c, d, e, f))
    uu, o, q, r = (np.zeros_like(f, order="F") if x is None else _fcol(x) for x in (uu, o, q, r))
    m1, m2 = f.shape
    _check(lib.lica_ptdiag1a(_dptr(a), _dptr(b), _dptr(c), _dptr(d), _dptr(e), _dptr(f), _dptr(uu), m1, m2, i1, i_n,
                             j1, jn, _dptr(o), _dptr(q), _dptr(r)), "lica_ptdiag1a")
    return uu, o, q, r


def poisson(a, b, c, g, mode=_abi.POISSON_THOMAS, rtol=1.0e-10, max_cycles=100, lib=None):
    """Batched POISSON tridiagonal systems (1d.f:3535-3564); arrays (nsys, n)."""
    lib = lib if lib is not None else _abi.load()
    a, b, c, g = (np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64) for x in (a, b, c, g))
    nsys, n = g.shape
    p = np.zeros_like(g)
    cyc = ctypes.c_int64()
    _check(lib.lica_poisson(nsys, n, _dptr(a), _dptr(b), _dptr(c), _dptr(g), _dptr(p), int(mode),
                            float(rtol), int(max_cycles), ctypes.byref(cyc)), "lica_poisson")
    return p
