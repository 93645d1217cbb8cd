"""GPU checks of the warm-start network and of the learned Poisson mode.

The MFMA kernels (BF16 operands, FP32 accumulation) are compared with the
NumPy restatement in net_ref.py, which quantises at the same points and
accumulates exactly: the forward output and the gradient agree to within
the accumulation-order differences of FP32 (a rare BF16 rounding flip of one
activation included).  The learned mode only chooses the initial guess of the
mixed-precision V-cycles, so its results meet the same backward-error
tolerance as the other multigrid modes.
"""
import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd.learned import Net, collect_pairs
from oracle import Oracle
import net_ref as R

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _feats(rng, lens):
    return [rng.uniform(-1.0, 1.0, (n, 4)).astype(np.float32) for n in lens]


def test_forward_matches_bf16_restatement():
    """Ragged sequences (shorter than a tile, a whole tile, several tiles)."""
    _say("net forward: lens 5, 64, 300, 1")
    rng = np.random.default_rng(3)
    net = Net(seed=5)
    p = net.params()
    lens = [5, 64, 300, 1]
    feats = _feats(rng, lens)
    y = net.forward(feats)
    yr, _ = R.forward(p, feats, quant=True)
    for a, b in zip(y, yr):
        assert a.shape == b.shape
    ya, yb = np.concatenate(y), np.concatenate(yr)
    scale = np.max(np.abs(yb))
    assert scale > 0
    err = np.abs(ya - yb)
    assert np.max(err) <= 1e-2 * scale, (np.max(err), scale)
    assert np.median(err) <= 1e-4 * scale, (np.median(err), scale)


def test_gradient_matches_bf16_restatement():
    _say("net gradient: lens 70, 130")
    rng = np.random.default_rng(4)
    net = Net(seed=6)
    p = net.params()
    lens = [70, 130]
    feats = _feats(rng, lens)
    targets = [rng.standard_normal(n).astype(np.float32) * 0.1 for n in lens]
    loss, g = net.grad(feats, targets)
    lr_, gr = R.loss_grad(p, feats, targets, quant=True)
    assert abs(loss - lr_) <= 1e-4 * lr_, (loss, lr_)
    o = 0
    for l in range(R.LAYERS):
        for n in (R.TAPS * R.COUT[l] * R.C, R.COUT[l]):
            a, b = g[o:o + n], gr[o:o + n]
            o += n
            den = np.linalg.norm(b)
            assert den > 0
            assert np.linalg.norm(a - b) <= 2e-2 * den, (l, n, np.linalg.norm(a - b) / den)
    assert o == len(g)


def test_parameters_round_trip():
    net = Net(seed=1)
    p = net.params()
    q = (p + np.float32(0.25)).astype(np.float32)
    net.set_params(q)
    assert np.array_equal(net.params(), q)


def test_training_fits_a_learnable_map():
    """A target that a convolution of the features represents exactly: Adam
    on minibatches brings the loss far below its starting value."""
    _say("net training on a synthetic linear target")
    rng = np.random.default_rng(7)
    nwin, win = 256, 256
    x = rng.uniform(-1.0, 1.0, (nwin, win, 4)).astype(np.float32)
    k = np.array([0.25, 0.5, 0.25])
    t = np.stack([0.5 * np.convolve(x[i, :, 0], k, mode="same") - 0.3 * x[i, :, 3] for i in range(nwin)])
    t = t.astype(np.float32)
    net = Net(seed=2)
    net.set_data(x, t)
    l0, _ = net.grad(list(x[:32]), list(t[:32]))
    l1 = net.train(400, batch=32, lr=1e-3, seed=1)
    _say(f"loss {l0:.4e} -> {l1:.4e}")
    assert l1 <= 0.05 * l0, (l0, l1)


def test_learned_mode_meets_tolerance():
    """Train on the solver's own POISSON pairs of the large case, then run the
    learned mode: every solve meets the backward-error tolerance and the
    interface stays within 1e-6 relative of the FP64 oracle."""
    case = L.single_large_case(1)
    _say("learned mode: collect pairs, train, 6 steps of the large case")
    fx, tx = collect_pairs([case], steps=6, per_step=16, mode=L.POISSON_MIXED)
    assert fx.shape[0] == 6 * 16
    net = Net(seed=0)
    net.set_data(fx, tx)
    net.train(100, batch=32, lr=1e-3)
    b, o = L.Batch([case]), Oracle(case)
    b.set_net(net)
    b.set_poisson(L.POISSON_LEARNED, 1e-10, 100)
    b.run(6)
    assert o.step(6) == 0
    s = b.scalars()
    assert s["POISSON_SOLVES"] == 18
    assert 0 < s["POISSON_RESID"] <= 1e-10
    zo = o.scalar_array()[3]
    assert abs(s["PLACE"] - zo) <= 1e-6 * zo
