"""ctypes binding of include/lica_mi355x.h (the C ABI of the library)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblica_mi355x.so")

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
P_dbl = ctypes.POINTER(ctypes.c_double)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_flt = ctypes.POINTER(ctypes.c_float)

# status codes (lica_mi355x.h)
OK = 0
BAND_OVERFLOW = 1
PSIF_BUG = 2
GRID_TOO_LARGE = 3
EXTRAP_NOT_CONVERGED = 4
BAND_CAPACITY = 5
NONFINITE = 6
WINDOW = 7
BAD_ARG = 20
UNSUPPORTED_DECK = 21
IO = 22
HIP_ERROR = 30
NO_DEVICE = 31
NOT_CONVERGED = 32

POISSON_THOMAS, POISSON_DIRECT, POISSON_MG64, POISSON_MIXED, POISSON_LEARNED, POISSON_PROJECTED = range(6)

FIELDS = {"W": 0, "P": 1, "T": 2, "PSI_CN": 3, "PSI_ZN": 4, "PSIF": 5, "PHI": 6,
          "SFGAS": 7, "WO": 8, "BAND": 9, "ALPHI": 10, "MFLVS": 11,
          "PA": 12, "PB": 13, "PC": 14, "PG": 15, "PPREV": 16, "PSOL": 17, "PRHS": 18}
SCALARS = ["TIME", "DT", "STEP", "PLACE", "DVMAX", "CFLM", "VOL_ORI", "VOL_IPHS",
           "QWALL", "STATUS", "SUBCYCLE_WARN", "MASS_WARN", "EXTRAP_HALVINGS",
           "EXTRAP_ITERS", "REINIT_CALLS", "BANDGEN_CALLS", "MASSCOR_ITERS",
           "POISSON_CYCLES", "POISSON_SOLVES", "POISSON_RESID", "DTCONST", "PSHIFT"]
NSCAL = 24


class CaseDesc(ctypes.Structure):
    """lica_case_desc: the deck values one boiling case depends on."""
    _fields_ = [("n", c_i64), ("r", c_i64), ("zl", c_dbl),
                ("re", c_dbl), ("visr", c_dbl), ("denr", c_dbl),
                ("surf_j", c_dbl), ("fr", c_dbl),
                ("prm", c_dbl), ("scr", c_dbl), ("tcr", c_dbl),
                ("tsat", c_dbl), ("ja", c_dbl),
                ("dt0", c_dbl), ("cflmax", c_dbl),
                ("idtopt", c_i64),
                ("z0", c_dbl), ("tslope", c_dbl)]

    def copy(self, **kw):
        c = CaseDesc()
        ctypes.pointer(c)[0] = self
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LicaError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        super().__init__(f"{where}: status {status} ({status_string(status)})")


_lib = None


def _declare(lib):
    V = ctypes.c_void_p
    PV = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "lica_parse_deck": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(CaseDesc)]),
        "lica_batch_create": (ctypes.c_int, [c_i64, ctypes.POINTER(CaseDesc), PV]),
        "lica_batch_init": (ctypes.c_int, [V]),
        "lica_batch_run": (ctypes.c_int, [V, c_i64]),
        "lica_batch_set_poisson": (ctypes.c_int, [V, c_i64, c_dbl, c_i64]),
        "lica_batch_set_block": (ctypes.c_int, [V, c_i64]),
        "lica_batch_use_graph": (ctypes.c_int, [V, c_i64]),
        "lica_batch_set_timing": (ctypes.c_int, [V, c_i64]),
        "lica_batch_sync": (ctypes.c_int, [V]),
        "lica_batch_get_field": (ctypes.c_int, [V, c_i64, c_i64, P_dbl, c_i64, P_i64]),
        "lica_batch_get_scalars": (ctypes.c_int, [V, c_i64, P_dbl]),
        "lica_batch_status": (ctypes.c_int, [V, P_i64]),
        "lica_batch_size": (c_i64, [V]),
        "lica_batch_get_progress": (ctypes.c_int, [V, P_i64, P_i64]),
        "lica_batch_last_timings": (ctypes.c_int, [V, P_dbl, c_i64]),
        "lica_batch_destroy": (None, [V]),
        "lica_batch_save": (ctypes.c_int, [V, ctypes.c_char_p]),
        "lica_batch_load": (ctypes.c_int, [ctypes.c_char_p, PV]),
        "lica_batch_write_fld": (ctypes.c_int, [V, c_i64, ctypes.c_char_p, c_i64]),
        "lica_init": (ctypes.c_int, [PV, ctypes.c_char_p, ctypes.c_char_p]),
        "lica_step": (ctypes.c_int, [V, c_i64]),
        "lica_determine_dt": (ctypes.c_int, [V, P_dbl, P_dbl]),
        "lica_rhsnlhs": (ctypes.c_int, [V, c_i64, P_dbl, P_dbl, P_dbl, P_dbl]),
        "lica_convrge1": (ctypes.c_int, [V, P_dbl]),
        "lica_transport": (ctypes.c_int, [V]),
        "lica_grid_coupling": (ctypes.c_int, [V, c_i64, P_dbl, P_dbl]),
        "lica_end_step": (ctypes.c_int, [V]),
        "lica_poisson": (ctypes.c_int, [c_i64, c_i64, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, c_i64, c_dbl, c_i64, P_i64]),
        # seams with the reference's argument lists: (3, 3, N3+1) Fortran-order arrays
        "lica_f_determine_dt": (ctypes.c_int, [V, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]),
        "lica_f_rhsnlhs": (ctypes.c_int, [V, P_i64, P_i64] + [P_dbl] * 14),
        "lica_f_convrge1": (ctypes.c_int, [V, P_dbl, P_dbl, P_dbl, P_dbl]),
        "lica_f_transport": (ctypes.c_int, [V, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]),
        "lica_f_grid_coupling": (ctypes.c_int, [V, P_i64, P_dbl, P_dbl, P_dbl, P_dbl]),
        "lica_f_poisson": (ctypes.c_int, [V] + [P_dbl] * 7),
        "lica_f_temp_grad": (ctypes.c_int, [V, P_dbl]),
        "lica_f_state": (ctypes.c_int, [V] + [P_dbl] * 8),
        # device addresses in and out, a hipStream_t (or None)
        "lica_poisson_noalloc": (ctypes.c_int, [c_i64, c_i64, V, V, V, V, V, c_i64, c_dbl, c_i64, P_i64, V]),
        "lica_trdiag1": (ctypes.c_int, [P_dbl] * 5 + [c_i64] * 6),
        "lica_trdiag2": (ctypes.c_int, [P_dbl] * 5 + [c_i64] * 6),
        "lica_trdiag3": (ctypes.c_int, [P_dbl] * 5 + [c_i64] * 6),
        "lica_trdiag1p": (ctypes.c_int, [P_dbl] * 4 + [c_i64] * 6),
        "lica_trdiag2p": (ctypes.c_int, [P_dbl] * 4 + [c_i64] * 6),
        "lica_trdiag3p": (ctypes.c_int, [P_dbl] * 4 + [c_i64] * 6),
        "lica_ptdiag1a": (ctypes.c_int, [P_dbl] * 7 + [c_i64] * 6 + [P_dbl] * 3),
        "lica_trdiag_lines_dev": (ctypes.c_int, [V] * 6 + [c_i64] * 7 + [V]),
        "lica_trdiag_periodic_lines_dev": (ctypes.c_int, [V] * 5 + [c_i64] * 6 + [V]),
        "lica_ptdiag_lines_dev": (ctypes.c_int, [V] * 7 + [c_i64] * 6 + [V] * 3 + [V]),
        "lica_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "lica_batch_set_net": (ctypes.c_int, [V, V]),
        "lica_batch_set_capture": (ctypes.c_int, [V, c_i64]),
        "lica_batch_set_mg_lds": (ctypes.c_int, [V, c_i64]),
        "lica_batch_warm_info": (ctypes.c_int, [V, P_dbl, c_i64]),
        "lica_batch_set_lanes": (ctypes.c_int, [V, c_i64]),
        "lica_batch_set_lane_split": (ctypes.c_int, [V, c_i64]),
        "lica_batch_read_fld": (ctypes.c_int, [V, c_i64, ctypes.c_char_p, c_i64]),
        "lica_batch_autotune": (ctypes.c_int, [V, c_i64]),
        "lica_batch_get_tuning": (ctypes.c_int, [V, P_i64, c_i64, P_i64]),
        "lica_net_param_count": (c_i64, []),
        "lica_net_create": (ctypes.c_int, [ctypes.c_uint64, PV]),
        "lica_net_destroy": (None, [V]),
        "lica_net_get_params": (ctypes.c_int, [V, P_flt]),
        "lica_net_set_params": (ctypes.c_int, [V, P_flt]),
        "lica_net_forward": (ctypes.c_int, [V, P_flt, P_i64, c_i64, P_flt]),
        "lica_net_grad": (ctypes.c_int, [V, P_flt, P_flt, P_i64, c_i64, P_flt, P_dbl]),
        "lica_net_set_data": (ctypes.c_int, [V, P_flt, P_flt, c_i64, c_i64]),
        "lica_net_train": (ctypes.c_int, [V, c_i64, c_i64, c_dbl, ctypes.c_uint64, P_dbl]),
        "lica_net_add_pairs": (ctypes.c_int, [V, V, c_i64, c_i64, ctypes.c_uint64, P_i64]),
        "lica_net_data_size": (ctypes.c_int, [V, P_i64, P_i64]),
        "lica_net_get_data": (ctypes.c_int, [V, P_flt, P_flt]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(path=None):
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if path is not None:
        return _declare(ctypes.CDLL(path))
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m lica_boiling1d_amd.build` "
                "(the boiling path has no CPU implementation)")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


# statuses a single case can end with (the reference's STOPs and a Poisson
# solve that hits its cycle cap); 20..31 are errors of the call itself
CASE_STATUSES = frozenset(range(1, 20)) | {32}


def is_case_status(status):
    """True for a status that a case of a batch ends with (not a call error)."""
    return int(status) in CASE_STATUSES


def status_string(status):
    try:
        return load().lica_status_string(int(status)).decode()
    except Exception:
        return "?"
