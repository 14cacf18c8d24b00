"""L-FTS driver on the hip platform: the device-resident saddle point and Langevin
update against the CPU oracle, against the rows the reference's own example run
recorded (examples/fts/Lamella.py:94-108), and the checkpoint/continue semantics
of lfts.py:569-629."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from oracle.lfts_oracle import ABLangevinOracle
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def diblock_params(nx, lx, model, ds, chi_n, f, nbar, dt, am, tol, max_iter, max_step, rec_dir, verbose=1):
    return {
        "nx": nx, "lx": lx, "chain_model": model, "ds": ds,
        "segment_lengths": {"A": 1.0, "B": 1.0},
        "chi_n": {"A,B": chi_n},
        "distinct_polymers": [{"volume_fraction": 1.0,
                               "blocks": [{"type": "A", "length": f}, {"type": "B", "length": 1 - f}]}],
        "langevin": {"max_step": max_step, "dt": dt, "nbar": nbar},
        "recording": {"dir": rec_dir, "recording_period": 1000, "sf_computing_period": 10,
                      "sf_recording_period": 10000},
        "saddle": {"max_iter": max_iter, "tolerance": tol},
        "am": {"max_hist": am[0], "start_error": am[1], "mix_min": am[2], "mix_init": am[3]},
        "verbose_level": verbose,
    }


@pytest.fixture(scope="module")
def lfts():
    from langevin_fts_amd import lfts as m
    return m


@pytest.mark.parametrize("nx,lx,am,tol,n_steps", [
    ([12, 10, 8], [3.0, 2.6, 2.2], (20, 1e-1, 0.1, 0.1), 1e-6, 3),
    # C5 itself: 64^3, L=6.0, the reference's L-FTS mixing (20, 0.5, 0.01, 0.01), tol 1e-4
    ([64, 64, 64], [6.0, 6.0, 6.0], (20, 5e-1, 0.01, 0.01), 1e-4, 2)])
def test_saddle_and_langevin_match_oracle(lfts, tmp_path, nx, lx, am, tol, n_steps):
    """C5's chain (discrete N=90, f=0.5, chiN=17, nbar=1e4, dt=0.8): every saddle point
    and Langevin step of the device path against the oracle with the same noise."""
    say(f"lfts vs oracle {nx}")
    chi_n, dt, nbar = 17.0, 0.8, 1e4
    max_iter = 200
    params = diblock_params(nx, lx, "discrete", 1 / 90, chi_n, 0.5, nbar, dt, am, tol, max_iter, 3,
                            str(tmp_path), verbose=0)
    sim = lfts.LFTS(params, random_seed=3)
    dev = sim.saddle_device
    M = int(np.prod(nx))
    plan = golden("plans")["diblock_f0.5_ds1_90|discrete|1"]
    inp = plan["input"]
    og = O.Grid(nx, lx)
    oracle = ABLangevinOracle(O.Solver(og, plan, inp["ds"], inp["bonds"]), O.AndersonMixing(M, *am),
                              chi_n, [(1.0, 1.0)], tol, max_iter)
    rng = np.random.default_rng(9)
    w_aux = np.stack([rng.standard_normal(M) * 2.0, rng.standard_normal(M) * 0.3])
    dev.set_fields(w_aux)
    w_ref = w_aux.copy()
    sigma = lfts.calculate_sigma(nbar, dt, M, np.prod(lx))
    noise_prev = np.zeros(M)
    dev.set_noise_prev(noise_prev[None, :])
    for step in range(n_steps + 1):
        if step > 0:
            noise = rng.normal(0.0, sigma, [1, M])
            dev.langevin_update(noise, dt)
            oracle.langevin_step(w_ref, phi_ref, noise_prev, noise[0], dt)
            noise_prev = noise[0]
            # a saddle point stopped at a 1e-4 error carries the evaluations' rounding
            # differences (1e-14) through 20+ Anderson steps: fields agree to about 1e-9
            np.testing.assert_allclose(dev.get_fields(), w_ref, rtol=0, atol=1e-8)
        n_iter, err, H, _ = dev.find_saddle_point(tol, max_iter, 0)
        phi_ref, H_ref, n_ref, err_ref = oracle.find_saddle_point(w_ref)
        say(f"  step {step}: iterations {n_iter} / oracle {n_ref}, error {err:.6e} / {err_ref:.6e}")
        assert n_iter == n_ref
        assert abs(err / err_ref - 1) < 1e-6
        assert abs(H - H_ref) < 1e-9
        np.testing.assert_allclose(dev.get_fields(), w_ref, rtol=0, atol=1e-8)
        np.testing.assert_allclose(dev.get_concentration(0), phi_ref["A"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(dev.get_concentration(1), phi_ref["B"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(dev.get_noise_prev()[0], noise_prev, rtol=0, atol=0)


def _parse_rows(text):
    num = r"([-+]?\d+\.?\d*(?:[eE][-+]?\d+)?)"
    rows = []
    for line in text.split("\n"):
        m = re.match(r"\s+(\d+)\s+" + num + r"\s+\[\s*" + num + r"\s*\]\s+" + num + r"\s+\[\s*" + num + r"\s*\]",
                     line)
        if m:
            rows.append([int(m.group(1))] + [float(m.group(i)) for i in range(2, 6)])
    return rows


def test_lamella_recorded_run(lfts, tmp_path, capsys):
    """examples/fts/Lamella.py with its own parameters and seeds: initial saddle
    point and Langevin steps 1-5 as the reference recorded them."""
    say("lamella recorded rows")
    rec = golden("drivers")["lamella_lfts"]["rows"]
    params = diblock_params([32, 32, 32], [8.0, 8.0, 8.0], "continuous", 1 / 16, 20.0, 0.5, 1024, 8.0,
                            (20, 5e-1, 0.01, 0.01), 1e-4, 100, 5, str(tmp_path))
    np.random.seed(12345)
    w_A = np.random.normal(0.0, 1.0, params["nx"])
    w_B = np.random.normal(0.0, 1.0, params["nx"])
    sim = lfts.LFTS(params=params, random_seed=12345)
    sim.run(initial_fields={"A": w_A, "B": w_B})
    out = capsys.readouterr().out
    rows = _parse_rows(out)
    with capsys.disabled():
        for r, ref in zip(rows, rec):
            say(f"  {r}  recorded {[ref['iter'], ref['mass_error'], ref['Q'], ref['H'], ref['error']]}")
    assert len(rows) == len(rec)
    for r, ref in zip(rows, rec):
        assert r[0] == ref["iter"]
        assert abs(r[1]) < 1e-12
        assert abs(r[2] / ref["Q"] - 1) < 1e-6
        assert abs(r[3] - ref["H"]) < 2e-8
        assert abs(r[4] / ref["error"] - 1) < 1e-5


def test_checkpoint_continue_run(lfts, tmp_path):
    """fields_%06d.mat holds w, phi, the PCG64 state and normal_noise_prev; a run
    continued from step 2 reaches the same step-4 fields as an uninterrupted run."""
    say("lfts checkpoint")
    from scipy.io import loadmat

    def params(d):
        p = diblock_params([16, 12, 10], [3.2, 2.8, 2.4], "discrete", 1 / 20, 15.0, 0.4, 1e3, 0.5,
                           (20, 1e-1, 0.1, 0.1), 1e-5, 200, 4, str(d), verbose=0)
        p["recording"]["recording_period"] = 2
        return p

    rng = np.random.default_rng(4)
    init = {"A": rng.standard_normal([16, 12, 10]), "B": rng.standard_normal([16, 12, 10])}
    a = tmp_path / "a"
    sim = lfts.LFTS(params(a), random_seed=21)
    sim.run(initial_fields=init)
    ref = loadmat(os.path.join(a, "fields_000004.mat"), squeeze_me=True)
    b = tmp_path / "b"
    os.makedirs(b)
    sim2 = lfts.LFTS(params(b), random_seed=999)
    sim2.continue_run(os.path.join(a, "fields_000002.mat"))
    got = loadmat(os.path.join(b, "fields_000004.mat"), squeeze_me=True)
    for key in ["w_A", "w_B", "phi_A", "phi_B", "normal_noise_prev"]:
        np.testing.assert_allclose(got[key], ref[key], rtol=0, atol=1e-10)
    assert got["random_state_state"] == ref["random_state_state"]
    assert int(got["langevin_step"]) == 4
