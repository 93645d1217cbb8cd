"""GPU check of lica_poisson_noalloc, the allocation-free variant of
lica_poisson for callers that cannot allocate device memory: the same
systems on device arrays give lica_poisson's solutions bit for bit (the same
kernels; the workspace is the library's fixed per-GPU array), on two calls in
a row with the second smaller than the first, in every mode; requests beyond
the fixed workspace are refused with LICA_TOO_LARGE before any work."""
import ctypes

import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd import _abi

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _systems(nsys, n, seed):
    """POISSON sign pattern (a, c > 0, b = -(a + c), the last row's
    coupling above the top face kept in its diagonal) with a density jump."""
    rng = np.random.default_rng(seed)
    rho = np.where(np.arange(n + 1)[None, :] < rng.integers(n // 4, 3 * n // 4, (nsys, 1)), 1.0, 1605.36)
    coef = (1.0 / rho) * rng.uniform(0.9, 1.1, (nsys, n + 1))
    a = coef[:, :-1].copy()
    c = coef[:, 1:].copy()
    a[:, 0] = 0.0
    b = -(a + c)
    g = rng.standard_normal((nsys, n))
    return a, b, c, g


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the HIP runtime the library itself uses
    (the test allocates the device arrays; the entry point under test does not)."""
    H2D, D2H = 1, 2

    def __init__(self):
        import os
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.rt = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(arr.nbytes)
        assert self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, self.H2D) == 0
        return p

    def get(self, p, shape):
        out = np.zeros(shape)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self, *ps):
        for p in ps:
            assert self.rt.hipFree(p) == 0


def _noalloc(a, b, c, g, mode, rtol=1e-10, max_cycles=100):
    lib = _abi.load()
    hip = _Hip()
    dev = [hip.put(x) for x in (a, b, c, g)]
    p = hip.alloc(g.nbytes)
    cyc = ctypes.c_int64(-1)
    nsys, n = g.shape
    rc = lib.lica_poisson_noalloc(nsys, n, *dev, p, int(mode), float(rtol), int(max_cycles), ctypes.byref(cyc), None)
    out = hip.get(p, g.shape)
    hip.free(p, *dev)
    return rc, out, cyc.value


@pytest.mark.parametrize("mode", [L.POISSON_THOMAS, L.POISSON_DIRECT, L.POISSON_MG64, L.POISSON_MIXED])
def test_noalloc_equals_lica_poisson(mode):
    for nsys, n, seed in ((37, 3000, 1), (5, 700, 2)):
        _say(f"lica_poisson_noalloc mode {mode}: {nsys} systems of {n} rows")
        a, b, c, g = _systems(nsys, n, seed)
        ref = L.poisson(a, b, c, g, mode=mode, rtol=1e-10, max_cycles=100)
        rc, p, cyc = _noalloc(a, b, c, g, mode)
        assert rc == 0, _abi.status_string(rc)
        assert np.array_equal(p, ref)
        if mode >= L.POISSON_MG64:
            assert 1 <= cyc <= 100
        else:
            assert cyc == 0


def test_noalloc_refuses_beyond_workspace():
    _say("lica_poisson_noalloc: requests beyond the fixed workspace")
    lib = _abi.load()
    hip = _Hip()
    for mode, nsys, n in ((L.POISSON_MG64, 200, 4096), (L.POISSON_THOMAS, 600, 4096)):
        x = hip.alloc(nsys * n * 8)
        cyc = ctypes.c_int64(0)
        rc = lib.lica_poisson_noalloc(nsys, n, x, x, x, x, x, int(mode), 1e-10, 100, ctypes.byref(cyc), None)
        hip.free(x)
        assert rc == 23, rc          # LICA_TOO_LARGE
