"""The NumPy restatement of the warm-start network: its gradient is pinned by
finite differences (unquantised), and its forward pass by a direct
convolution written independently (np.convolve per channel)."""
import numpy as np

import net_ref as R


def small_params(rng, scale=0.1):
    n = sum(R.TAPS * c * R.C + c for c in R.COUT)
    return (rng.standard_normal(n) * scale).astype(np.float32)


def test_forward_matches_direct_convolution():
    rng = np.random.default_rng(0)
    p = small_params(rng)
    f = rng.standard_normal((40, 4))
    y, _ = R.forward(p, [f], quant=False)
    layers = R.unpack(p)
    x = np.zeros((40, R.C)); x[:, :4] = f
    for l, (w, b) in enumerate(layers):
        z = np.zeros((40, R.COUT[l]))
        for co in range(R.COUT[l]):
            for ci in range(R.C):
                # correlation with taps t: x[cell + t - HALF]
                z[:, co] += np.convolve(x[:, ci], w[::-1, co, ci], mode="same")
            z[:, co] += b[co]
        if l < R.LAYERS - 1:
            x = np.zeros((40, R.C)); x[:, :R.COUT[l]] = np.maximum(z, 0.0)
    assert np.allclose(y[0], z[:, 0], rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = small_params(rng, 0.2).astype(np.float64)
    feats = [rng.standard_normal((30, 4)), rng.standard_normal((25, 4))]
    targets = [rng.standard_normal(30), rng.standard_normal(25)]
    loss, g = R.loss_grad(p, feats, targets, quant=False)
    idx = rng.choice(len(p), 40, replace=False)
    # include last-layer and first-layer weights and biases
    idx = np.concatenate([idx, [0, len(p) - 1, len(p) - 17]])
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e; pm[i] -= e
        lp, _ = R.loss_grad(pp, feats, targets, quant=False)
        lm, _ = R.loss_grad(pm, feats, targets, quant=False)
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39], dtype=np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5
