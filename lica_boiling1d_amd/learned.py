"""Learned warm start of the Poisson solve.

The POISSON solution of a substep (1d.f:3529-3543) is, to about 1e-6, a
combination of the case's solutions at the same substep in the last steps:
``LICA_POISSON_PROJECTED`` starts the mixed-precision V-cycles from B, the
combination of those two and the last solution that minimises |g - A B|_2.
What is left, p - B, sits at the interface.  ``LICA_POISSON_LEARNED`` adds a
1-D convolutional network's correction (4 layers of 21 taps, 32 channels,
ReLU; BF16 weights and activations, FP32 accumulation on MFMA; about 46k
parameters) that maps, cell by cell,

    r / max|r|,  a / max|b|,  b / max|b|,  c / max|b|      (r = g - A B)

to a direction d over a window of 256 rows around the interface, where the
error p - B left by the projection is largest (a and c jump with the
density there); the residual fixes its length, guess = B + alpha d with
alpha = <A d, r> / <A d, A d>, so a useless direction costs nothing.  Cases
whose B already meets the tolerance skip the network.  The V-cycles always
continue to the same backward-error tolerance, so multigrid decides the
answer and the network only saves cycles.  Training pairs come
from the solver itself (``collect_pairs``): a batch runs with capture on and
after every step windows of every case's captured POISSON system and
starting point go straight into the network's data set in device memory.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import LicaError

NFEAT = 4
WINDOW = 256


def _check(rc, where):
    if rc != 0:
        raise LicaError(rc, where)


def _f(a):
    return a.ctypes.data_as(_abi.P_flt)


class Net:
    """The warm-start network (parameters on the GPU)."""

    def __init__(self, seed=0, lib=None):
        self._lib = lib if lib is not None else _abi.load()
        self._h = ctypes.c_void_p()
        _check(self._lib.lica_net_create(int(seed), ctypes.byref(self._h)), "lica_net_create")
        self.nparam = int(self._lib.lica_net_param_count())

    def close(self):
        if self._h:
            self._lib.lica_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        _check(self._lib.lica_net_get_params(self._h, _f(out)), "lica_net_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _check(self._lib.lica_net_set_params(self._h, _f(p)), "lica_net_set_params")

    def forward(self, feats):
        """feats: list of (len, 4) arrays -> list of (len,) outputs"""
        lens = np.array([len(f) for f in feats], dtype=np.int64)
        x = np.ascontiguousarray(np.concatenate(feats).astype(np.float32))
        y = np.zeros(int(lens.sum()), dtype=np.float32)
        _check(self._lib.lica_net_forward(self._h, _f(x), lens.ctypes.data_as(_abi.P_i64), len(lens), _f(y)),
               "lica_net_forward")
        return np.split(y, np.cumsum(lens)[:-1])

    def grad(self, feats, targets):
        lens = np.array([len(f) for f in feats], dtype=np.int64)
        x = np.ascontiguousarray(np.concatenate(feats).astype(np.float32))
        t = np.ascontiguousarray(np.concatenate(targets).astype(np.float32))
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = ctypes.c_double()
        _check(self._lib.lica_net_grad(self._h, _f(x), _f(t), lens.ctypes.data_as(_abi.P_i64), len(lens), _f(g),
                                       ctypes.byref(loss)), "lica_net_grad")
        return loss.value, g

    def add_pairs(self, batch, windows_per_case=4, window=WINDOW, seed=0):
        """Append windows of the batch's captured POISSON systems to the data
        set (device to device); returns the number of windows added."""
        added = ctypes.c_int64()
        _check(self._lib.lica_net_add_pairs(self._h, batch._h, int(windows_per_case), int(window), int(seed),
                                            ctypes.byref(added)), "lica_net_add_pairs")
        return added.value

    def data(self):
        """The data set: features (n, len, 4) as stored (BF16 values), targets (n, len)."""
        n, ln = ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.lica_net_data_size(self._h, ctypes.byref(n), ctypes.byref(ln)), "lica_net_data_size")
        x = np.zeros((n.value, ln.value, NFEAT), dtype=np.float32)
        t = np.zeros((n.value, ln.value), dtype=np.float32)
        if n.value:
            _check(self._lib.lica_net_get_data(self._h, _f(x), _f(t)), "lica_net_get_data")
        return x, t

    def set_data(self, feats, targets):
        """feats (n, WINDOW, 4), targets (n, WINDOW); replaces the data set"""
        n, ln = targets.shape
        x = np.ascontiguousarray(feats.reshape(-1, NFEAT).astype(np.float32))
        t = np.ascontiguousarray(targets.reshape(-1).astype(np.float32))
        _check(self._lib.lica_net_set_data(self._h, _f(x), _f(t), n, ln), "lica_net_set_data")

    def train(self, iters, batch=128, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        _check(self._lib.lica_net_train(self._h, int(iters), int(batch), float(lr), int(seed), ctypes.byref(loss)),
               "lica_net_train")
        return loss.value


def features(a, b, c, g, base):
    """Per-cell features of one POISSON system (rows 1..n) and starting
    point; the arithmetic of the device feature kernels, before the BF16
    rounding."""
    ab = a * np.r_[0.0, base[:-1]]
    cb = c * np.r_[base[1:], 0.0]
    r = g - ((b * base + ab) + cb)
    sr = np.max(np.abs(r)) or 1.0
    sb = np.max(np.abs(b)) or 1.0
    f = np.stack([r / sr, a / sb, b / sb, c / sb], axis=1)
    return f.astype(np.float32)


def target(p, base):
    """Training target of a window: p - base scaled to unit maximum (zero
    where they agree)."""
    e = p - base
    se = np.max(np.abs(e)) or 1.0
    return (e / se).astype(np.float32)


def collect_pairs(net, cases, steps, windows_per_case=4, window=WINDOW, seed=0, mode=None, rtol=1e-10, lib=None,
                  warmup=2):
    """Run the cases in the projected warm-start mode (or `mode`) with
    capture on, `warmup` steps to fill the solution history, then `steps`
    steps after each of which windows of every healthy case's captured
    POISSON system (substeps 1, 2, 3 in turn) join the network's data set.
    Returns the number of windows added."""
    from .solver import Batch
    from . import _abi as A
    b = Batch(cases, lib=lib if lib is not None else net._lib)
    b.set_poisson(A.POISSON_PROJECTED if mode is None else mode, rtol, 200)
    if warmup:
        b.run_all(warmup)
    total = 0
    for s in range(steps):
        b.set_capture(True, substep=1 + s % 3)
        b.run_all(1)
        total += net.add_pairs(b, windows_per_case, window, seed * 1000003 + s)
    b.close()
    return total
