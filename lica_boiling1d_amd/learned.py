"""Learned warm start of the Poisson solve.

A 1-D convolutional network (4 layers of 21 taps, 32 channels, ReLU; BF16
weights and activations, FP32 accumulation on MFMA; about 46k parameters)
maps, cell by cell, the features of a POISSON system (1d.f:3529-3543)

    g / max|g|,  a / max|b|,  c / max|b|,  P / max|P|

(g the right-hand side, a/b/c the three diagonals, P the pressure held before
the solve, i.e. the previous substep's) to the correction (p - P) / max|P| of
the solution p.  ``LICA_POISSON_LEARNED`` starts the mixed-precision V-cycles
from P + max|P| * network; the V-cycles always continue to the same residual
tolerance, so multigrid decides the answer and the network only saves
cycles.  Training pairs come from the solver itself (``collect_pairs``): a
batch runs with capture on and keeps, for every case, P before and after the
last substep's solve.
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

    def set_data(self, feats, targets):
        """feats (n, WINDOW, 4), targets (n, WINDOW)"""
        n, ln = targets.shape
        x = np.ascontiguousarray(feats.reshape(-1, NFEAT).astype(np.float32))
        t = np.ascontiguousarray(targets.reshape(-1).astype(np.float32))
        _check(self._lib.lica_net_set_data(self._h, _f(x), _f(t), n, ln), "lica_net_set_data")

    def train(self, iters, batch=128, lr=1e-3, seed=0):
        loss = ctypes.c_double()
        _check(self._lib.lica_net_train(self._h, int(iters), int(batch), float(lr), int(seed), ctypes.byref(loss)),
               "lica_net_train")
        return loss.value


def features(a, b, c, g, p):
    """Per-cell features and the solution scale of one POISSON system (rows
    1..n); the same arithmetic as the device feature kernel."""
    sg = np.max(np.abs(g)) or 1.0
    sb = np.max(np.abs(b)) or 1.0
    sp = np.max(np.abs(p)) or 1.0
    f = np.stack([g / sg, a / sb, c / sb, p / sp], axis=1)
    return f.astype(np.float32), sp


def collect_pairs(cases, steps, per_step=4, window=WINDOW, seed=0, mode=None, lib=None):
    """Run the cases with capture on and cut (features, target) windows of
    `window` cells from every case's last-substep POISSON system after every
    step.  Returns (feats (m, window, 4), targets (m, window))."""
    from .solver import Batch
    from . import _abi as A
    rng = np.random.default_rng(seed)
    b = Batch(cases, lib=lib)
    b.set_poisson(A.POISSON_DIRECT if mode is None else mode, 1e-10, 200)
    _check(b._lib.lica_batch_set_capture(b._h, 1), "set_capture")
    fx, tx = [], []
    for _ in range(steps):
        b.run_all(1)
        st, _ = b.progress()
        for i, cs in enumerate(cases):
            if st[i] != 0 or int(cs.n) < window:
                continue
            a, bb, c, g = (b.field(k, i) for k in ("PA", "PB", "PC", "PRHS"))
            p0, p1 = b.field("PPREV", i), b.field("PSOL", i)
            f, sp = features(a, bb, c, g, p0)
            t = ((p1 - p0) / sp).astype(np.float32)
            for _ in range(per_step):
                o = int(rng.integers(0, int(cs.n) - window + 1))
                fx.append(f[o:o + window])
                tx.append(t[o:o + window])
    b.close()
    return np.array(fx), np.array(tx)
