"""Learned warm start of the Poisson solve.

A 1-D convolutional network (4 layers of 21 taps, 32 channels, ReLU; BF16
weights and activations, FP32 accumulation on MFMA; about 46k parameters)
maps, cell by cell, the features of a POISSON system (1d.f:3529-3543)

    g / max|g|,  a / max|b|,  c / max|b|,  Q / max|Q|

(g the right-hand side, a/b/c the three diagonals, Q the previous substep's
solution: the pressure held before the solve plus the p(1) that UCALC
subtracted from it, 1d.f:3433-3470) to the correction (p - Q) / max|Q| of the
solution p.  ``LICA_POISSON_LEARNED`` starts the mixed-precision V-cycles from
Q + max|Q| * network; the V-cycles always continue to the same residual
tolerance, so multigrid decides the answer and the network only saves
cycles.  Training pairs come from the solver itself (``collect_pairs``): a
batch runs with capture on in a multigrid mode, and after every step windows
of every case's last POISSON system go from its capture rows straight into
the network's data set in device memory.
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


def features(a, b, c, g, q):
    """Per-cell features and the solution scale of one POISSON system (rows
    1..n, q the previous solution); the same arithmetic as the device
    feature kernels, before the BF16 rounding."""
    sg = np.max(np.abs(g)) or 1.0
    sb = np.max(np.abs(b)) or 1.0
    sq = np.max(np.abs(q)) or 1.0
    f = np.stack([g / sg, a / sb, c / sb, q / sq], axis=1)
    return f.astype(np.float32), sq


def collect_pairs(net, cases, steps, windows_per_case=4, window=WINDOW, seed=0, mode=None, rtol=1e-10, lib=None):
    """Run the cases `steps` steps in a multigrid mode (mixed precision by
    default) with capture on and append, after every step, windows of every
    healthy case's last POISSON system to the network's data set.  Returns the
    number of windows added."""
    from .solver import Batch
    from . import _abi as A
    b = Batch(cases, lib=lib if lib is not None else net._lib)
    b.set_poisson(A.POISSON_MIXED if mode is None else mode, rtol, 200)
    b.set_capture(True)
    total = 0
    for s in range(steps):
        b.run_all(1)
        total += net.add_pairs(b, windows_per_case, window, seed * 1000003 + s)
    b.close()
    return total
