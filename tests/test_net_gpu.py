"""GPU checks of the warm-start network and of the learned Poisson mode.

The MFMA kernels (BF16 operands, FP32 accumulation) are compared with the
NumPy restatement in net_ref.py, which quantises at the same points and
accumulates exactly: the forward output and the gradient agree to within
the accumulation-order differences of FP32 (a rare BF16 rounding flip of one
activation included).  The projected and learned modes only choose the
initial guess of the mixed-precision V-cycles, so their results meet the same
backward-error tolerance as the other multigrid modes.
"""
import numpy as np
import pytest

import lica_boiling1d_amd as L
from lica_boiling1d_amd.learned import Net, collect_pairs, features, target
from oracle import Oracle
from parity_util import ulp_sensitivity
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


@pytest.mark.parametrize("mode", [L.POISSON_MIXED, L.POISSON_PROJECTED])
def test_device_pairs_equal_host_features(mode):
    """Windows cut on the device from the capture rows equal the features and
    targets computed on the host from the same rows (window = whole column,
    so its offset is 0); the starting point is the last solution in the
    mixed mode and the projection base in the projected mode."""
    _say(f"device training pairs, 3 cases n=128, mode {mode}")
    cases = [L.deck_case(128, 2).copy(z0=z) for z in (0.05, 0.0607, 0.08)]
    net = Net(seed=0)
    b = L.Batch(cases)
    b.set_poisson(mode, 1e-10, 100)
    b.run(3)
    b.set_capture(True, substep=2)
    b.run(1)
    assert net.add_pairs(b, 1, 128, seed=0) == 3
    x, t = net.data()
    assert x.shape == (3, 128, 4) and t.shape == (3, 128)
    for i in range(3):
        a, bb, c, g = (b.field(k, i) for k in ("PA", "PB", "PC", "PRHS"))
        q, p = b.field("PPREV", i), b.field("PSOL", i)
        f = features(a, bb, c, g, q)
        assert np.array_equal(x[i], R.bf16(f).astype(np.float32)), i
        assert np.array_equal(t[i], target(p, q)), i
    # appending keeps what is held; the window length is fixed by the data set
    b.run(1)
    assert net.add_pairs(b, 2, 128, seed=1) == 6
    x2, t2 = net.data()
    assert x2.shape == (9, 128, 4)
    assert np.array_equal(x2[:3], x) and np.array_equal(t2[:3], t)
    with pytest.raises(L.LicaError):
        net.add_pairs(b, 1, 64)
    b.close()


def _base_scales(a, b, c, g, x):
    """Backward error of x and the feature scales s_r, s_b of one system
    (rows 1..n), in the residual's order of operations on the device."""
    v = b * x
    v[1:] = v[1:] + a[1:] * x[:-1]
    v[:-1] = v[:-1] + c[:-1] * x[1:]
    r = g - v
    ak = np.r_[0.0, np.abs(a[1:])]
    ck = np.r_[np.abs(c[:-1]), 0.0]
    m0, m1 = np.max(np.abs(r)), np.max(np.abs(b))
    m2, m3, m4 = np.max((ak + np.abs(b)) + ck), np.max(np.abs(x)), np.max(np.abs(g))
    den = m2 * m3 + m4
    return (m0 / den if den > 0 else m0), (m0 or 1.0), (m1 or 1.0)


def _start_test(a, b, c, g, x, rtol):
    """The multigrid start test on x: backward error with the residual
    accumulated as ((g - b x) - a x-) - c x+, and whether x is taken as is."""
    r = g - b * x
    r[1:] = r[1:] - a[1:] * x[:-1]
    r[:-1] = r[:-1] - c[:-1] * x[1:]
    ak = np.r_[0.0, np.abs(a[1:])]
    ck = np.r_[np.abs(c[:-1]), 0.0]
    rm, fm = np.max(np.abs(r)), np.max(np.abs(g))
    den = np.max((ak + np.abs(b)) + ck) * np.max(np.abs(x)) + fm
    res = rm / den if den > 0 else rm
    return res, (not rm > fm) and (not res > rtol)


@pytest.mark.parametrize("mode", [L.POISSON_LEARNED])
def test_learned_base_scales_equal_host(mode):
    """The warm-start base kernel's pass over B also yields the backward
    error and the scales the network's features use, and the multigrid
    solve's start test on B; they equal, bit for bit, the ones computed on
    the host from the captured rows of the same substep (the base B before
    the network's correction, the matrix, the RHS)."""
    _say(f"warm-start base scales, 3 cases n=200, mode {mode}")
    cases = [L.deck_case(200, 2).copy(z0=z) for z in (0.05, 0.0607, 0.08)]
    net = Net(seed=0)
    b = L.Batch(cases)
    b.set_net(net)
    b.set_poisson(mode, 1e-10, 100)
    b.run(3)
    b.set_capture(True, substep=3)
    b.run(1)
    info = b.warm_info()
    for i in range(3):
        a, bb, c, g = (b.field(k, i) for k in ("PA", "PB", "PC", "PRHS"))
        x = b.field("PPREV", i)
        assert tuple(info[i, 1:4]) == _base_scales(a, bb, c, g, x), i
        if mode == L.POISSON_LEARNED:
            assert info[i, 0] == 1.0   # the window covers the whole column
        res, take = _start_test(a, bb, c, g, x, 1e-10)
        assert info[i, 5] == res, i
        # the network clears the verdict of the cases whose guess it moved
        moved = mode == L.POISSON_LEARNED and info[i, 1] > 1e-10
        assert info[i, 6] == (1.0 if take and not moved else 0.0), i
    b.close()
    net.close()


def test_learned_correction_matches_layer_kernels():
    """The per-case kernel of the learned mode (features, the four layers in
    LDS, the step alpha) against the layer-by-layer training-path kernels:
    the host computes the features of the captured base, runs them through
    lica_net_forward and fits alpha = <A d, r> / <A d, A d> in FP64; the
    device's alpha agrees to summation order."""
    _say("learned correction vs layer kernels, 3 cases n=200")
    cases = [L.deck_case(200, 2).copy(z0=z) for z in (0.05, 0.0607, 0.08)]
    net = Net(seed=0)
    assert collect_pairs(net, cases, steps=3, windows_per_case=2, window=128) > 0
    net.train(50, batch=4, lr=1e-3)
    b = L.Batch(cases)
    b.set_net(net)
    b.set_poisson(L.POISSON_LEARNED, 1e-10, 100)
    b.run(2)
    b.set_capture(True, substep=3)
    b.run(1)
    info = b.warm_info()
    applied = 0
    for i in range(3):
        a, bb, c, g = (b.field(k, i) for k in ("PA", "PB", "PC", "PRHS"))
        x = b.field("PPREV", i)
        if not info[i, 1] > 1e-10:
            assert info[i, 4] == 0.0
            continue
        applied += 1
        d = net.forward([features(a, bb, c, g, x)])[0].astype(np.float64)
        ad = bb * d
        ad[1:] += a[1:] * d[:-1]
        ad[:-1] += c[:-1] * d[1:]
        r = g - ((bb * x + np.r_[0.0, a[1:] * x[:-1]]) + np.r_[c[:-1] * x[1:], 0.0])
        alpha = np.dot(ad, r) / np.dot(ad, ad)
        assert abs(info[i, 4] - alpha) <= 1e-9 * abs(alpha), (i, info[i, 4], alpha)
    assert applied > 0
    b.close()
    net.close()


def test_add_pairs_needs_capture_and_multigrid():
    cases = [L.deck_case(128, 2)]
    net = Net(seed=0)
    b = L.Batch(cases)
    b.set_poisson(L.POISSON_MIXED, 1e-10, 100)
    b.run(1)
    with pytest.raises(L.LicaError):
        net.add_pairs(b, 1, 64)
    b.set_poisson(L.POISSON_DIRECT)
    b.set_capture(True)
    b.run(1)
    with pytest.raises(L.LicaError):
        net.add_pairs(b, 1, 64)


def test_learned_two_lanes_equal_one():
    """With the cases split into two or four sets on as many streams, the per-case
    network kernel reads each lane's case order: the fields and the warm-start
    info equal the single sequence's bit for bit (stream launches and graph
    replay)."""
    cases = L.batched_cases(6, seed=7, nmin=128, nmax=512)
    _say("learned mode, 2 and 4 lanes vs one, 6 cases, 2 x 3 steps")
    net = Net(seed=0)
    runs = []
    for lanes, graph in ((1, False), (2, False), (2, True), (4, False)):
        b = L.Batch(cases)
        b.set_net(net)
        b.set_poisson(L.POISSON_LEARNED, 1e-10, 100)
        b.set_lanes(lanes)
        b.use_graph(graph)
        runs.append(b)
    for _ in range(2):
        for b in runs:
            b.run(3)
        ref_info = runs[0].warm_info()
        for b in runs[1:]:
            assert np.array_equal(ref_info, b.warm_info())
        for i in range(len(cases)):
            for f in ("W", "P", "T", "PHI"):
                ref = runs[0].field(f, i)
                for b in runs[1:]:
                    assert np.array_equal(ref, b.field(f, i)), (f, i)
    for b in runs:
        b.close()
    net.close()


def _cycles(b, steps):
    s0 = b.scalars()
    b.run(steps)
    s = b.scalars()
    return (s["POISSON_CYCLES"] - s0["POISSON_CYCLES"]) / (s["POISSON_SOLVES"] - s0["POISSON_SOLVES"]), s


def test_projected_mode_large_case():
    """History projection on the large case (N=8192, r=4): every solve meets
    the stopping test, the interface stays within 1e-10 relative of the FP64
    oracle, and after the start-up transient the V-cycles per solve fall
    below the zero start's."""
    case = L.single_large_case(1)
    _say("projected warm start, large case, 30 steps")
    o = Oracle(case)
    assert o.step(30) == 0
    cyc = {}
    for mode in (L.POISSON_MIXED, L.POISSON_PROJECTED):
        b = L.Batch([case])
        b.set_poisson(mode, 1e-10, 100)
        b.run(20)
        cyc[mode], s = _cycles(b, 10)
        assert 0 < s["POISSON_RESID"] <= 1e-10
        zo = o.scalar_array()[3]
        assert abs(s["PLACE"] - zo) <= 1e-10 * zo
    _say(f"cycles/solve: zero start {cyc[L.POISSON_MIXED]:.2f}, projected {cyc[L.POISSON_PROJECTED]:.2f}")
    assert cyc[L.POISSON_PROJECTED] < cyc[L.POISSON_MIXED]


def test_learned_mode_meets_tolerance():
    """Train on the solver's own POISSON pairs of the large case, then run the
    learned mode: every solve meets the stopping test and the interface
    stays within 1e-10 relative of the FP64 oracle."""
    case = L.single_large_case(1)
    _say("learned mode: collect pairs, train, 8 steps of the large case")
    net = Net(seed=0)
    assert collect_pairs(net, [case], steps=6, windows_per_case=16) == 6 * 16
    net.train(100, batch=32, lr=1e-3)
    b, o = L.Batch([case]), Oracle(case)
    b.set_net(net)
    b.set_poisson(L.POISSON_LEARNED, 1e-10, 100)
    b.run(8)
    assert o.step(8) == 0
    s = b.scalars()
    assert s["POISSON_SOLVES"] == 24
    assert 0 < s["POISSON_RESID"] <= 1e-10
    zo = o.scalar_array()[3]
    assert abs(s["PLACE"] - zo) <= 1e-10 * zo


@pytest.mark.parametrize("which,steps", [("deck", 100), ("large", 20)])
def test_field_errors_against_oracle(which, steps):
    """Per-field maximum relative error against the FP64 oracle for the plain
    FP64 V-cycles, the mixed-precision V-cycles and the projected warm start
    (printed).  Every field within 1e-10 relative or 10x the oracle's own
    change under a one-ulp perturbation of its input (the direct mode's
    bar), the interface position and the wall heat flux within 1e-10."""
    case = L.deck_case() if which == "deck" else L.single_large_case(1)
    o = Oracle(case)
    assert o.step(steps) == 0
    fields = ("W", "P", "T", "PHI")
    ref = {f: o.field(f) for f in fields}
    so = dict(zip(L.SCALARS, o.scalar_array()))
    sens = ulp_sensitivity(case, steps, fields)
    _say(f"{which} {steps} steps one-ulp sensitivity: " + " ".join(f"{k}={v:.1e}" for k, v in sens.items()))
    for mode, name in ((L.POISSON_MG64, "mg64"), (L.POISSON_MIXED, "mixed"), (L.POISSON_PROJECTED, "projected")):
        b = L.Batch([case])
        b.set_poisson(mode, 1e-10, 100)
        b.run(steps)
        s = b.scalars()
        err = {f: float(np.max(np.abs(b.field(f) - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300)) for f in ref}
        err["PLACE"] = abs(s["PLACE"] - so["PLACE"]) / abs(so["PLACE"])
        err["QWALL"] = abs(s["QWALL"] - so["QWALL"]) / abs(so["QWALL"])
        _say(f"{which} {steps} steps {name}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
        assert err["PLACE"] <= 1e-10 and err["QWALL"] <= 1e-10, err
        for f in fields:
            assert err[f] <= max(1e-10, 10.0 * sens[f]), (f, err[f], sens[f])
        b.close()
