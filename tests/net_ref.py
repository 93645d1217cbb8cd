"""NumPy restatement of the warm-start network (tests only): the same
layers, layouts and BF16 quantisation points as the MFMA kernels, with
exact (float64) accumulation.  quant=False gives the unquantised network,
whose gradient is checked by finite differences."""
import numpy as np

C, TAPS, HALF, LAYERS = 32, 21, 10, 4
COUT = (32, 32, 32, 16)


def bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(params):
    """flat parameters -> [(W[tap][co][ci], b[co])]"""
    out, o = [], 0
    for l in range(LAYERS):
        n = TAPS * COUT[l] * C
        w = params[o:o + n].astype(np.float64).reshape(TAPS, COUT[l], C)
        o += n
        b = params[o:o + COUT[l]].astype(np.float64)
        o += COUT[l]
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers]).astype(np.float64)


def conv(x, w):
    """x (len, C), w (TAPS, cout, C) -> (len, cout), zero padding"""
    n = x.shape[0]
    xp = np.zeros((n + 2 * HALF, C))
    xp[HALF:HALF + n] = x
    y = np.zeros((n, w.shape[1]))
    for t in range(TAPS):
        y += xp[t:t + n] @ w[t].T
    return y


def forward(params, feats, quant=True):
    """feats: list of (len, 4) -> (outputs, cache)"""
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    outs, caches = [], []
    for f in feats:
        x = np.zeros((len(f), C))
        x[:, :4] = f
        x = q(x)
        acts = [x]
        for l, (w, b) in enumerate(layers):
            z = conv(acts[-1], q(w)) + b
            if l == LAYERS - 1:
                y = z[:, 0]
            else:
                a = q(np.maximum(z, 0.0))
                full = np.zeros((len(f), C))
                full[:, :COUT[l]] = a
                acts.append(full)
        outs.append(y)
        caches.append(acts)
    return outs, caches


def loss_grad(params, feats, targets, quant=True):
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    outs, caches = forward(params, feats, quant)
    count = sum(len(t) for t in targets)
    loss = sum(float(np.sum((y - t) ** 2)) for y, t in zip(outs, targets)) / count
    grads = [(np.zeros_like(w), np.zeros_like(b)) for w, b in layers]
    for y, t, acts in zip(outs, targets, caches):
        n = len(y)
        dz = np.zeros((n, C))
        dz[:, 0] = q(2.0 * (y - t) / count)
        for l in range(LAYERS - 1, -1, -1):
            w, _ = layers[l]
            x = acts[l]
            xp = np.zeros((n + 2 * HALF, C))
            xp[HALF:HALF + n] = x
            gw, gb = grads[l]
            for tp in range(TAPS):
                gw[tp] += dz[:, :COUT[l]].T @ xp[tp:tp + n]
            gb += dz[:, :COUT[l]].sum(axis=0)
            if l > 0:
                # dX = conv(dZ, W flipped in taps, transposed)
                wt = np.zeros((TAPS, C, C))
                for tp in range(TAPS):
                    wt[tp, :, :COUT[l]] = q(w[TAPS - 1 - tp]).T
                dzp = np.zeros((n, C))
                dzp[:, :COUT[l]] = dz[:, :COUT[l]]
                dx = q(conv(dzp, wt))
                dz = dx * (acts[l] > 0.0)
    return loss, pack(grads)
