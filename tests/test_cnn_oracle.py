"""The numpy restatement of the learned-guess network (oracle/cnn_oracle.py)
pinned by arithmetic it must satisfy: bf16 rounding of known values, the forward
pass against an explicit loop over taps and channels, and every analytic
gradient block against central finite differences of the exact (unrounded)
network. CPU only."""
import numpy as np
import pytest

from oracle import cnn_oracle as K


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-5, 1.0 + 2.0 ** -8], dtype=np.float32)
    # 1 + 2^-9 is a tie between 1 and 1 + 2^-7: even mantissa (1.0); 1 + 3*2^-9 rounds up
    want = [1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, None, 1.0]
    got = K.bf16(vals)
    for g, w, v in zip(got, want, vals):
        if w is not None:
            assert g == w
        assert abs(g - v) <= abs(v) * 2.0 ** -8


@pytest.mark.parametrize("n_in", [1, 3])
def test_forward_matches_explicit_loops(n_in):
    rng = np.random.default_rng(0)
    p = K.init_params(1, n_in).astype(np.float64)
    w0o, b0o, who, _, _, _ = K.layout(n_in)
    p[b0o:who] = rng.normal(0, 0.1, K.C)
    x = rng.standard_normal((n_in, 4, 5, 6))
    y, acts = K.forward(p, x if n_in > 1 else x[0], rounding=False)
    w0, b0, hid, wl, bl = K.unpack(p, n_in)
    v = (1, 4, 0)
    # first layer at one voxel, every input channel
    z0 = b0.copy()
    for c in range(n_in):
        for t, (dx, dy, dz) in enumerate(K.OFFSETS):
            z0 += w0[c, t] * x[c, (v[0] + dx) % 4, (v[1] + dy) % 5, (v[2] + dz) % 6]
    np.testing.assert_allclose(acts[0][v], np.maximum(z0, 0), rtol=1e-13, atol=1e-13)
    # one hidden layer at one voxel
    w, b = hid[2]
    h = acts[2]
    z = b.copy()
    for t, (dx, dy, dz) in enumerate(K.OFFSETS):
        for co in range(K.C):
            for ci in range(K.C):
                z[co] += w[t, co, ci] * h[(v[0] + dx) % 4, (v[1] + dy) % 5, (v[2] + dz) % 6, ci]
    np.testing.assert_allclose(acts[3][v], np.maximum(z, 0), rtol=1e-12, atol=1e-12)
    # output layer
    out = bl
    for t, (dx, dy, dz) in enumerate(K.OFFSETS):
        out += np.dot(wl[t], acts[4][(v[0] + dx) % 4, (v[1] + dy) % 5, (v[2] + dz) % 6])
    assert abs(out - y[v]) < 1e-12


@pytest.mark.parametrize("block,n_in", [("W0", 1), ("b0", 1), ("W1", 1), ("b1", 1), ("W4", 1), ("b4", 1),
                                        ("WL", 1), ("bL", 1), ("W0", 3), ("b0", 3), ("W1", 3), ("WL", 3)])
def test_gradients_match_finite_differences(block, n_in):
    rng = np.random.default_rng(2)
    w0o, b0o, who, wlo, blo, _ = K.layout(n_in)
    p = K.init_params(3, n_in).astype(np.float64)
    p[b0o:who] = 0.05
    for l in range(K.NHID):
        p[who + l * K.HSZ + K.TAPS * K.C * K.C:who + (l + 1) * K.HSZ] = 0.05
    x = rng.standard_normal((n_in, 4, 4, 3))
    target = rng.standard_normal((4, 4, 3))
    _, g = K.loss_and_gradients(p, x, target, rounding=False)
    lo = {"W0": w0o, "b0": b0o, "WL": wlo, "bL": blo}
    hid = {"W1": (0, 0), "b1": (0, 1), "W4": (3, 0), "b4": (3, 1)}
    if block in lo:
        start = lo[block]
        size = {"W0": n_in * K.TAPS * K.C, "b0": K.C, "WL": K.TAPS * K.C, "bL": 1}[block]
    else:
        l, part = hid[block]
        start = who + l * K.HSZ + (K.TAPS * K.C * K.C if part else 0)
        size = K.C if part else K.TAPS * K.C * K.C
    idx = start + rng.choice(size, size=min(6, size), replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (K.loss_and_gradients(pp, x, target, rounding=False)[0] -
              K.loss_and_gradients(pm, x, target, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (block, i, fd, g[i])
