"""Periodic 3D CNN of the learned saddle-point guess, restated in numpy (test
infrastructure only: the checker of langevin_fts_amd's fts_cnn kernels).

conv(1->32) ReLU, 4 x [conv(32->32) ReLU], conv(32->1); 3x3x3 kernels with circular
padding. The arithmetic is FP64; values are rounded to bf16 (round to nearest even)
where the device keeps them in bf16: hidden weights, hidden activations after
ReLU, and the activation gradients that feed the hidden layers' products.

Parameter layout (FP32 master vector): W0[27][32], b0[32], then per hidden layer
W[27][co][ci], b[32], then WL[27][32], bL[1]. Tap t = 9(dx+1) + 3(dy+1) + (dz+1);
tap t reads the input at voxel + (dx, dy, dz), periodically.
"""
import numpy as np

C = 32
TAPS = 27
NHID = 4
HSZ = TAPS * C * C + C
P_W0, P_B0 = 0, TAPS * C
P_WH = P_B0 + C
P_WL = P_WH + NHID * HSZ
P_BL = P_WL + TAPS * C
N_PARAMS = P_BL + 1
OFFSETS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def bf16(a):
    """Round float32 values to bf16 (nearest, ties to even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shifted(x, off):
    """x at voxel + off (periodic), for x of shape (nx, ny, nz[, c])."""
    return np.roll(x, shift=(-off[0], -off[1], -off[2]), axis=(0, 1, 2))


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w0 = p[P_W0:P_B0].reshape(TAPS, C)
    b0 = p[P_B0:P_WH]
    hid = []
    for l in range(NHID):
        base = P_WH + l * HSZ
        hid.append((p[base:base + TAPS * C * C].reshape(TAPS, C, C), p[base + TAPS * C * C:base + HSZ]))
    wl = p[P_WL:P_BL].reshape(TAPS, C)
    bl = p[P_BL]
    return w0, b0, hid, wl, bl


def init_params(seed):
    """He-normal weights, zero biases (the driver's initialisation)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    p[P_W0:P_B0] = rng.normal(0.0, np.sqrt(2.0 / TAPS), TAPS * C)
    for l in range(NHID):
        base = P_WH + l * HSZ
        p[base:base + TAPS * C * C] = rng.normal(0.0, np.sqrt(2.0 / (TAPS * C)), TAPS * C * C)
    p[P_WL:P_BL] = rng.normal(0.0, np.sqrt(1.0 / (TAPS * C)), TAPS * C)
    return p


def _q(a, rounding):
    return bf16(a) if rounding else np.asarray(a, dtype=np.float64)


def forward(p, x, rounding=True):
    """x: (nx, ny, nz) input in network units. Returns y and the activations.
    rounding=False keeps every value in FP64 (the exact network, for derivative checks)."""
    w0, b0, hid, wl, bl = unpack(p)
    x = np.asarray(x, dtype=np.float32).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)
    z = b0 + sum(shifted(x, OFFSETS[t])[..., None] * w0[t] for t in range(TAPS))
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
    w0, b0, hid, wl, bl = unpack(p)
    if rounding:
        x = np.asarray(x, dtype=np.float32).astype(np.float64)
        target = np.asarray(target, dtype=np.float32).astype(np.float64)
    y, acts = forward(p, x, rounding)
    M = y.size
    d = y - target
    loss = np.mean(d * d)
    dy = 2.0 * d / M
    g = np.zeros(N_PARAMS)
    h = acts[-1]
    g[P_WL:P_BL] = np.stack([np.einsum("xyz,xyzc->c", dy, shifted(h, OFFSETS[t])) for t in range(TAPS)]).ravel()
    g[P_BL] = dy.sum()
    # dL/d(pre-activation of the last hidden layer)
    dh = sum(shifted(dy, tuple(-o for o in OFFSETS[t]))[..., None] * wl[t] for t in range(TAPS))
    dz = _q(np.where(h > 0, dh, 0.0), rounding)
    for l in range(NHID - 1, -1, -1):
        w, _ = hid[l]
        xin = acts[l]
        base = P_WH + l * HSZ
        g[base:base + TAPS * C * C] = np.stack(
            [np.einsum("xyzo,xyzc->oc", dz, shifted(xin, OFFSETS[t])) for t in range(TAPS)]).ravel()
        g[base + TAPS * C * C:base + HSZ] = dz.sum(axis=(0, 1, 2))
        wq = _q(w, rounding)
        dx = sum(np.einsum("xyzo,oc->xyzc", shifted(dz, tuple(-o for o in OFFSETS[t])), wq[t]) for t in range(TAPS))
        dz = _q(np.where(xin > 0, dx, 0.0), rounding)
    g[P_W0:P_B0] = np.stack([np.einsum("xyzc,xyz->c", dz, shifted(x, OFFSETS[t])) for t in range(TAPS)]).ravel()
    g[P_B0:P_WH] = dz.sum(axis=(0, 1, 2))
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on FP64 copies (step counts from 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
