"""The line solvers of the 3-D code (1d.f:3721-4034) on the GPU: the
library's host entry points against the oracle bit for bit, and the
device-array forms against the host forms."""
import ctypes
import sys

import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd import _abi as abi
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tri_inputs(m1, m2, seed):
    g = np.random.default_rng(seed)
    a, c, r = (g.uniform(-1.0, 1.0, (m1, m2)) for _ in range(3))
    b = (3.0 + g.uniform(0.0, 1.0, (m1, m2))) * np.where(g.uniform(size=(m1, m2)) < 0.5, -1.0, 1.0)
    return a, b, c, r


def _penta_inputs(m1, m2, seed):
    g = np.random.default_rng(seed)
    b, c, d, e, f = (g.uniform(-1.0, 1.0, (m1, m2)) for _ in range(5))
    a = 5.0 + g.uniform(0.0, 1.0, (m1, m2))
    return a, b, c, d, e, f


@pytest.mark.parametrize("m1,m2,l1,l2,ll1,ll2", [(4096, 300, 1, 300, 1, 4096), (1000, 64, 5, 60, 3, 998),
                                                 (77, 2, 1, 2, 1, 77)])
def test_trdiag_bit_exact(m1, m2, l1, l2, ll1, ll2):
    """lica_trdiag1/2/3 on the GPU: UU and R bit for bit the oracle's."""
    _say(f"TRDIAG1/2/3 on the GPU, {m1} x {m2}, rows {l1}..{l2}, lines {ll1}..{ll2}")
    a, b, c, r = _tri_inputs(m1, m2, 11)
    for kind in (1, 3):
        uu, r_out = L.trdiag(kind, a, b, c, r, l1, l2, ll1, ll2)
        ref_uu, ref_r = O.trdiag_lines(a, b, c, r, l1, l2, ll1, ll2, overwrite_r=(kind == 3))
        assert np.array_equal(uu, ref_uu), kind
        assert np.array_equal(r_out, ref_r), kind


def test_periodic_and_pentadiagonal_bit_exact():
    """lica_trdiag2p and lica_ptdiag1a on the GPU: bit for bit the oracle's."""
    _say("TRDIAG2P and PTDIAG1A on the GPU, 2048 lines")
    a, b, c, f = _tri_inputs(2048, 128, 12)
    assert np.array_equal(L.trdiag_periodic(2, a, b, c, f, 1, 128, 1, 2048), O.trdiag_periodic_lines(a, b, c, f, 1, 128, 1, 2048))
    a, b, c, d, e, f = _penta_inputs(2048, 200, 13)
    out = L.ptdiag1a(a, b, c, d, e, f, 2, 2047, 1, 200)
    ref = O.ptdiag_lines(a, b, c, d, e, f, 2, 2047, 1, 200)
    for x, y in zip(out, ref):
        assert np.array_equal(x, y)


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the HIP runtime the library uses."""
    H2D, D2H = 1, 2

    def __init__(self):
        import os
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.rt = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.held = []

    def put(self, arr):
        """a device copy of a (m1, m2) array in the column-major layout"""
        arr = np.asfortranarray(arr, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, self.H2D) == 0
        self.held.append(p)
        return p

    def get(self, p, shape):
        out = np.zeros(shape, order="F")
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def close(self):
        for p in self.held:
            assert self.rt.hipFree(p) == 0
        self.held = []


def test_device_forms_equal_host_forms():
    """lica_trdiag_lines_dev / _periodic_lines_dev / lica_ptdiag_lines_dev on
    device arrays (column-major) give the host entry points' results bit for
    bit."""
    lib = abi.load()
    hip = _Hip()
    m1, m2 = 3000, 96
    _say(f"device-array forms of the line solvers, {m1} x {m2}")
    z = np.zeros((m1, m2))
    a, b, c, r = _tri_inputs(m1, m2, 14)
    da, db, dc, dr = (hip.put(x) for x in (a, b, c, r))
    du, dw = hip.put(z), hip.put(z)
    assert lib.lica_trdiag_lines_dev(da, db, dc, dr, du, dw, m1, m2, 2, m2, 1, m1, 1, None) == 0
    assert hip.rt.hipDeviceSynchronize() == 0
    uu, r3 = L.trdiag(3, a, b, c, r, 2, m2, 1, m1)
    assert np.array_equal(hip.get(du, (m1, m2)), uu) and np.array_equal(hip.get(dr, (m1, m2)), r3)

    df, dw3 = hip.put(r), hip.put(np.zeros((3 * m1, m2)))
    assert lib.lica_trdiag_periodic_lines_dev(da, db, dc, df, dw3, m1, m2, 1, m2, 1, m1, None) == 0
    assert hip.rt.hipDeviceSynchronize() == 0
    assert np.array_equal(hip.get(df, (m1, m2)), L.trdiag_periodic(1, a, b, c, r, 1, m2, 1, m1))

    pa = _penta_inputs(m1, m2, 15)
    dp = [hip.put(x) for x in pa]
    do = [hip.put(z) for _ in range(4)]
    assert lib.lica_ptdiag_lines_dev(*dp, do[0], m1, m2, 1, m1, 1, m2, do[1], do[2], do[3], None) == 0
    assert hip.rt.hipDeviceSynchronize() == 0
    for d, x in zip(do, L.ptdiag1a(*pa, 1, m1, 1, m2)):
        assert np.array_equal(hip.get(d, (m1, m2)), x)
    hip.close()
