"""Periodic 3D CNN of the learned saddle-point guess, restated in numpy (test
infrastructure only: the checker of langevin_fts_amd's fts_cnn kernels).

conv(n_in->32) ReLU, 4 x [conv(32->32) ReLU], conv(32->1); 3x3x3 kernels with
circular padding. The arithmetic is FP64; values are rounded to bf16 (round to
nearest even) where the device keeps them in bf16: hidden weights, hidden
activations after ReLU, and the activation gradients that feed the hidden layers'
products.

Parameter layout (FP32 master vector): W0[n_in][27][32], b0[32], then per hidden
layer W[27][co][ci], b[32], then WL[27][32], bL[1]. Tap t = 9(dx+1) + 3(dy+1) +
(dz+1); tap t reads the input at voxel + (dx, dy, dz), periodically. Inputs are
(nx, ny, nz) for one channel or (n_in, nx, ny, nz).
"""
import numpy as np

C = 32
TAPS = 27
NHID = 4
HSZ = TAPS * C * C + C
OFFSETS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def layout(n_in=1):
    """Offsets (w0, b0, wh, wl, bl, n) of the parameter vector."""
    w0 = 0
    b0 = w0 + n_in * TAPS * C
    wh = b0 + C
    wl = wh + NHID * HSZ
    bl = wl + TAPS * C
    return w0, b0, wh, wl, bl, bl + 1


P_W0, P_B0, P_WH, P_WL, P_BL, N_PARAMS = layout(1)


def n_params(n_in=1):
    return layout(n_in)[5]


def bf16(a):
    """Round float32 values to bf16 (nearest, ties to even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shifted(x, off):
    """x at voxel + off (periodic), for x of shape (nx, ny, nz[, c])."""
    return np.roll(x, shift=(-off[0], -off[1], -off[2]), axis=(0, 1, 2))


def unpack(p, n_in=1):
    p = np.asarray(p, dtype=np.float64)
    w0o, b0o, who, wlo, blo, _ = layout(n_in)
    w0 = p[w0o:b0o].reshape(n_in, TAPS, C)
    b0 = p[b0o:who]
    hid = []
    for l in range(NHID):
        base = who + l * HSZ
        hid.append((p[base:base + TAPS * C * C].reshape(TAPS, C, C), p[base + TAPS * C * C:base + HSZ]))
    wl = p[wlo:blo].reshape(TAPS, C)
    bl = p[blo]
    return w0, b0, hid, wl, bl


def init_params(seed, n_in=1):
    """He-normal weights, zero biases (the driver's initialisation)."""
    rng = np.random.default_rng(seed)
    w0o, b0o, who, wlo, blo, n = layout(n_in)
    p = np.zeros(n, dtype=np.float32)
    p[w0o:b0o] = rng.normal(0.0, np.sqrt(2.0 / (TAPS * n_in)), n_in * TAPS * C)
    for l in range(NHID):
        base = who + l * HSZ
        p[base:base + TAPS * C * C] = rng.normal(0.0, np.sqrt(2.0 / (TAPS * C)), TAPS * C * C)
    p[wlo:blo] = rng.normal(0.0, np.sqrt(1.0 / (TAPS * C)), TAPS * C)
    return p


def _channels(x):
    x = np.asarray(x)
    return x[None] if x.ndim == 3 else x


def _q(a, rounding):
    return bf16(a) if rounding else np.asarray(a, dtype=np.float64)


def forward(p, x, rounding=True):
    """x: input in network units, (nx, ny, nz) or (n_in, nx, ny, nz). Returns y and the
    activations. rounding=False keeps every value in FP64 (the exact network, for
    derivative checks)."""
    x = _channels(x)
    n_in = x.shape[0]
    w0, b0, hid, wl, bl = unpack(p, n_in)
    x = np.asarray(x, dtype=np.float32).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)
    z = b0 + sum(shifted(x[c], OFFSETS[t])[..., None] * w0[c, t] for c in range(n_in) for t in range(TAPS))
    acts = [_q(np.maximum(z, 0.0), rounding)]
    for w, b in hid:
        wq = _q(w, rounding)
        h = acts[-1]
        z = b + sum(np.einsum("xyzc,oc->xyzo", shifted(h, OFFSETS[t]), wq[t]) for t in range(TAPS))
        acts.append(_q(np.maximum(z, 0.0), rounding))
    h = acts[-1]
    y = bl + sum(np.einsum("xyzc,c->xyz", shifted(h, OFFSETS[t]), wl[t]) for t in range(TAPS))
    return y, acts


def loss_and_gradients(p, x, target, rounding=True):
    """Loss = mean (y - target)^2 and its gradient w.r.t. the parameter vector."""
    x = _channels(x)
    n_in = x.shape[0]
    w0o, b0o, who, wlo, blo, n = layout(n_in)
    w0, b0, hid, wl, bl = unpack(p, n_in)
    if rounding:
        x = np.asarray(x, dtype=np.float32).astype(np.float64)
        target = np.asarray(target, dtype=np.float32).astype(np.float64)
    y, acts = forward(p, x, rounding)
    M = y.size
    d = y - target
    loss = np.mean(d * d)
    dy = 2.0 * d / M
    g = np.zeros(n)
    h = acts[-1]
    g[wlo:blo] = np.stack([np.einsum("xyz,xyzc->c", dy, shifted(h, OFFSETS[t])) for t in range(TAPS)]).ravel()
    g[blo] = dy.sum()
    # dL/d(pre-activation of the last hidden layer)
    dh = sum(shifted(dy, tuple(-o for o in OFFSETS[t]))[..., None] * wl[t] for t in range(TAPS))
    dz = _q(np.where(h > 0, dh, 0.0), rounding)
    for l in range(NHID - 1, -1, -1):
        w, _ = hid[l]
        xin = acts[l]
        base = who + l * HSZ
        g[base:base + TAPS * C * C] = np.stack(
            [np.einsum("xyzo,xyzc->oc", dz, shifted(xin, OFFSETS[t])) for t in range(TAPS)]).ravel()
        g[base + TAPS * C * C:base + HSZ] = dz.sum(axis=(0, 1, 2))
        wq = _q(w, rounding)
        dx = sum(np.einsum("xyzo,oc->xyzc", shifted(dz, tuple(-o for o in OFFSETS[t])), wq[t]) for t in range(TAPS))
        dz = _q(np.where(xin > 0, dx, 0.0), rounding)
    g[w0o:b0o] = np.stack([np.einsum("xyzc,xyz->c", dz, shifted(x[c], OFFSETS[t]))
                           for c in range(n_in) for t in range(TAPS)]).ravel()
    g[b0o:who] = dz.sum(axis=(0, 1, 2))
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on FP64 copies (step counts from 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
