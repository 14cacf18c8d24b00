"""fts_grid_fft_backward_noalloc (no reference counterpart: the allocation-free form of
fts_grid_fft_backward for callers that cannot allocate) against fts_grid_fft_backward,
called through the C ABI: bit-identical results on two calls in a row on one GPU, the
second with a smaller input, and a clear error for a grid beyond its workspace."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def spectra(g, capi, batch, dt, seed):
    """batch forward transforms of random real fields, as the backward input"""
    a = np.random.default_rng(seed).standard_normal((batch, g.M)).astype(dt)
    x = capi.DeviceArray.from_host(a)
    spec = capi.DeviceArray(batch * g.Mc * 2 * np.dtype(dt).itemsize)
    g.fft_forward(x.ptr, spec.ptr, batch, capi.FTS_FP64 if dt == np.float64 else capi.FTS_FP32)
    return a, spec


def backward_both(ctx, capi, g, spec, batch, dt):
    p = capi.FTS_FP64 if dt == np.float64 else capi.FTS_FP32
    out_ref = capi.DeviceArray(batch * g.M * np.dtype(dt).itemsize)
    out_na = capi.DeviceArray(batch * g.M * np.dtype(dt).itemsize)
    g.fft_backward(spec.ptr, out_ref.ptr, batch, p)
    capi.check(capi.lib().fts_grid_fft_backward_noalloc(g.h, p, ctypes.c_void_p(spec.ptr),
                                                        ctypes.c_void_p(out_na.ptr), batch))
    ctx.synchronize()
    return out_ref.download((batch, g.M), dt), out_na.download((batch, g.M), dt)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_backward_equals_backward(prec):
    from langevin_fts_amd import capi

    dt = np.float64 if prec == "fp64" else np.float32
    ctx = capi.Context(0)
    # first call: 128^3 batch 3 (one spectrum per chunk of the workspace in FP64),
    # second call: a smaller grid and batch, on the same device workspace
    for shape, batch in [((128, 128, 128), 3), ((48, 40, 36), 2)]:
        say(f"noalloc backward {prec} {shape} batch {batch}")
        g = capi.Grid(ctx, list(shape), [1.0] * 3)
        a, spec = spectra(g, capi, batch, dt, seed=sum(shape) + batch)
        ref, na = backward_both(ctx, capi, g, spec, batch, dt)
        assert np.array_equal(ref, na)
        tol = 1e-12 if prec == "fp64" else 3e-6
        assert np.max(np.abs(na - a)) < tol * np.max(np.abs(a))


def test_noalloc_backward_refuses_grid_beyond_workspace():
    from langevin_fts_amd import capi

    ctx = capi.Context(0)
    say("noalloc backward: 512x128x64 FP64 (34.6 MB of spectrum) exceeds the workspace")
    g = capi.Grid(ctx, [512, 128, 64], [1.0] * 3)
    spec = capi.DeviceArray(g.Mc * 16)
    out = capi.DeviceArray(g.M * 8)
    rc = capi.lib().fts_grid_fft_backward_noalloc(g.h, capi.FTS_FP64, ctypes.c_void_p(spec.ptr),
                                                  ctypes.c_void_p(out.ptr), 1)
    assert rc != 0
    assert b"workspace" in capi.lib().fts_last_error()
