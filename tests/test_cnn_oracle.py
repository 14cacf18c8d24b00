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


def test_forward_matches_explicit_loops():
    rng = np.random.default_rng(0)
    p = K.init_params(1).astype(np.float64)
    p[K.P_B0:K.P_WH] = rng.normal(0, 0.1, K.C)
    x = rng.standard_normal((4, 5, 6))
    y, acts = K.forward(p, x, rounding=False)
    w0, b0, hid, wl, bl = K.unpack(p)
    v = (1, 4, 0)
    # first layer at one voxel
    z0 = b0.copy()
    for t, (dx, dy, dz) in enumerate(K.OFFSETS):
        z0 += w0[t] * x[(v[0] + dx) % 4, (v[1] + dy) % 5, (v[2] + dz) % 6]
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


@pytest.mark.parametrize("block", ["W0", "b0", "W1", "b1", "W4", "b4", "WL", "bL"])
def test_gradients_match_finite_differences(block):
    rng = np.random.default_rng(2)
    p = K.init_params(3).astype(np.float64)
    p[K.P_B0:K.P_WH] = 0.05
    for l in range(K.NHID):
        p[K.P_WH + l * K.HSZ + K.TAPS * K.C * K.C:K.P_WH + (l + 1) * K.HSZ] = 0.05
    x = rng.standard_normal((4, 4, 3))
    target = rng.standard_normal((4, 4, 3))
    _, g = K.loss_and_gradients(p, x, target, rounding=False)
    lo = {"W0": K.P_W0, "b0": K.P_B0, "WL": K.P_WL, "bL": K.P_BL}
    hid = {"W1": (0, 0), "b1": (0, 1), "W4": (3, 0), "b4": (3, 1)}
    if block in lo:
        start = lo[block]
        size = {"W0": K.TAPS * K.C, "b0": K.C, "WL": K.TAPS * K.C, "bL": 1}[block]
    else:
        l, part = hid[block]
        start = K.P_WH + l * K.HSZ + (K.TAPS * K.C * K.C if part else 0)
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
