"""FFT schedule autotuner of the N^3 wave-local kernels (fts_grid_get_fft_schedule /
fts_grid_set_fft_schedule): every launch-shape candidate gives the same propagators,
Q and concentrations to rounding (relative 1e-12 in FP64, 1e-5 in FP32), bit for bit
at 32^3 and 64^3 (where tuning is on by default; at the other sizes the defaults stay
unless FTS_FFT_TUNE=1), for discrete and continuous chains; the chosen shape is one of
the candidates, and results through a non-default schedule still meet the oracle's
1e-10 bar."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu

SLAB = {32: (128, 256), 48: (256, 384), 64: (256, 320, 512), 96: (512, 1024), 128: (512, 1024)}
PENCIL = (4, 8, 16)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def capi():
    from langevin_fts_amd import capi as c
    return c


@pytest.fixture(scope="module")
def ctx(capi):
    return capi.Context(0)


def evaluate(capi, g, precision, name, w):
    plan = golden("plans")[name]
    inp = plan["input"]
    s = capi.Solver(g, capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"]), precision)
    s.compute_statistics(w)
    key = plan["propagators"][0]["key"]
    n = plan["propagators"][0]["max_n_segment"]
    return (s.get_total_partition(0), s.get_total_concentration("A"), s.get_total_concentration("B"),
            s.propagator(key, n))


# the discrete step (slab half-step mode) and the continuous chain's forward transform with
# two spectra and Richardson step, the modes the autotuner times on their own
@pytest.mark.parametrize("name", ["diblock_f0.5_ds1_90|discrete|0", "diblock_f0.5_ds1_16|continuous|0"])
@pytest.mark.parametrize("n", [32, 48, 64, 96, 128])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_every_schedule_gives_identical_results(capi, ctx, n, prec, name):
    precision = capi.FTS_FP64 if prec == "fp64" else capi.FTS_FP32
    M = n ** 3
    rng = np.random.default_rng(n)
    w = {"A": rng.standard_normal(M), "B": rng.standard_normal(M)}
    g = capi.Grid(ctx, [n] * 3, [6.0] * 3)
    bd, nl, note = g.fft_schedule(precision)
    say(f"{n}^3 {prec} {name}: slab {bd} threads, pencil {nl} columns; {note}")
    assert bd in SLAB[n] and nl in PENCIL
    ref = evaluate(capi, g, precision, name, w)
    tol = 1e-12 if prec == "fp64" else 1e-5
    exact = n in (32, 64)
    for b in SLAB[n]:
        for c in PENCIL:
            g.set_fft_schedule(b, c, precision)
            assert g.fft_schedule(precision)[:2] == (b, c)
            out = evaluate(capi, g, precision, name, w)
            assert abs(out[0] / ref[0] - 1) < tol, (b, c)
            for a, r in zip(out[1:], ref[1:]):
                assert np.max(np.abs(a - r)) / np.max(np.abs(r)) < tol, (b, c)
                if exact:
                    assert np.array_equal(a, r), (b, c)


def test_tuned_schedule_meets_oracle(capi, ctx):
    """32^3 continuous diblock through a non-default schedule against the FP64 oracle"""
    name = "diblock_f0.36_ds0.01|continuous|0"
    plan = golden("plans")[name]
    inp = plan["input"]
    n, lx = 32, [3.3] * 3
    M = n ** 3
    rng = np.random.default_rng(5)
    w = {"A": rng.standard_normal(M), "B": rng.standard_normal(M)}
    g = capi.Grid(ctx, [n] * 3, lx)
    g.set_fft_schedule(128, 16)
    s = capi.Solver(g, capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"]), capi.FTS_FP64)
    s.compute_statistics(w)
    ref = O.Solver(O.Grid([n] * 3, lx), plan, inp["ds"], inp["bonds"])
    ref.compute_statistics(w)
    assert abs(s.get_total_partition(0) / ref.get_total_partition(0) - 1) < 1e-10
    for m in ("A", "B"):
        a, r = s.get_total_concentration(m), ref.get_total_concentration(m)
        assert np.max(np.abs(a - r)) / np.max(np.abs(r)) < 1e-10


def test_schedule_errors(capi, ctx):
    g = capi.Grid(ctx, [32] * 3, [4.0] * 3)
    with pytest.raises(RuntimeError, match="no slab kernel"):
        g.set_fft_schedule(320, 8)
    with pytest.raises(RuntimeError, match="pencil columns"):
        g.set_fft_schedule(256, 6)
    h = capi.Grid(ctx, [20, 18, 16], [4.0] * 3)
    assert h.fft_schedule()[:2] == (0, 0)
    with pytest.raises(RuntimeError, match="wave-local"):
        h.set_fft_schedule(256, 8)


def test_tuning_off_keeps_defaults(capi, ctx):
    """FTS_FFT_TUNE=0 at grid creation: the default shapes, no timing. The pencil pass
    defaults to 8 kz columns per wavefront, and to 4 in FP64 from 96^3."""
    import os

    defaults = {32: 256, 48: 384, 64: 320, 96: 1024, 128: 1024}
    os.environ["FTS_FFT_TUNE"] = "0"
    try:
        for n in (32, 64, 96, 128):
            g = capi.Grid(ctx, [n] * 3, [4.0] * 3)
            for precision in (capi.FTS_FP64, capi.FTS_FP32):
                bd, nl, note = g.fft_schedule(precision)
                assert (bd, nl) == (defaults[n], 4 if precision == capi.FTS_FP64 and n >= 96 else 8)
                assert note == "defaults"
    finally:
        os.environ.pop("FTS_FFT_TUNE", None)
