"""SCFT driver on the hip platform against the rows the reference's own example
recorded (examples/scft/Gyroid.py:91-96: 32^3 gyroid, continuous AB, f=0.36,
chiN=20, box relaxation), and the non-relaxing path against the oracle."""
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def gyroid_fields(nx):
    """Initial gyroid fields of examples/scft/Gyroid.py (double-gyroid level-set form)."""
    x = (np.arange(nx[0]) + 1) * 2 * np.pi / nx[0]
    y = (np.arange(nx[1]) + 1) * 2 * np.pi / nx[1]
    z = (np.arange(nx[2]) + 1) * 2 * np.pi / nx[2]
    xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
    c1 = np.sqrt(8.0 / 3.0) * (np.cos(xx) * np.sin(yy) * np.sin(2.0 * zz) + np.cos(yy) * np.sin(zz) * np.sin(2.0 * xx)
                               + np.cos(zz) * np.sin(xx) * np.sin(2.0 * yy))
    c2 = np.sqrt(4.0 / 3.0) * (np.cos(2.0 * xx) * np.cos(2.0 * yy) + np.cos(2.0 * yy) * np.cos(2.0 * zz)
                               + np.cos(2.0 * zz) * np.cos(2.0 * xx))
    return -0.3164 * c1 + 0.1074 * c2, 0.3164 * c1 - 0.1074 * c2


def params(nx, lx, box, max_iter, f=0.36, model="continuous", ds=1 / 100, chi_n=20.0):
    return {"nx": nx, "lx": lx, "box_is_altering": box, "chain_model": model, "ds": ds,
            "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": chi_n},
            "distinct_polymers": [{"volume_fraction": 1.0,
                                   "blocks": [{"type": "A", "length": f}, {"type": "B", "length": 1 - f}]}],
            "optimizer": {"name": "am", "max_hist": 20, "start_error": 1e-2, "mix_min": 0.1, "mix_init": 0.1},
            "max_iter": max_iter, "tolerance": 1e-8}


def test_gyroid_recorded_iterations(capsys):
    from langevin_fts_amd import scft

    say("scft gyroid recorded rows")
    rec = golden("drivers")["gyroid_scft"]["rows"]
    nx = [32, 32, 32]
    w_a, w_b = gyroid_fields(nx)
    calc = scft.SCFT(params(nx, [3.3, 3.3, 3.3], True, len(rec)))
    calc.run(initial_fields={"A": w_a, "B": w_b})
    out = capsys.readouterr().out
    num = r"([-+]?\d+\.?\d*(?:[eE][-+]?\d+)?)"
    rows = []
    for line in out.split("\n"):
        m = re.match(r"\s+(\d+)\s+" + num + r"\s+\[\s*" + num + r"\s*\]\s+" + num + r"\s+" + num +
                     r"\s+\[\s*" + num + r",\s*" + num + r",\s*" + num + r"\s*\]", line)
        if m:
            rows.append([int(m.group(1))] + [float(m.group(i)) for i in range(2, 9)])
    with capsys.disabled():
        for r, ref in zip(rows, rec):
            say(f"  {r}\n  recorded {[ref['iter'], ref['mass_error'], ref['Q'], ref['H'], ref['error']] + ref['lx']}")
    assert len(rows) == len(rec)
    for r, ref in zip(rows, rec):
        assert r[0] == ref["iter"]
        assert abs(r[1]) < 1e-12
        assert abs(r[2] / ref["Q"] - 1) < 1e-6
        assert abs(r[3] - ref["H"]) < 2e-9
        assert abs(r[4] / ref["error"] - 1) < 1e-6
        np.testing.assert_allclose(r[5:8], ref["lx"], rtol=0, atol=2e-7)


def test_scft_iterations_match_oracle(capsys):
    """Fixed box, discrete chain: each iteration's fields against the oracle solver
    with the reference's residual, error and Anderson mixing restated in numpy."""
    from langevin_fts_amd import scft

    nx, lx = [16, 12, 10], [3.2, 2.9, 2.4]
    say(f"scft vs oracle {nx}")
    p = params(nx, lx, False, 8, f=0.5, model="discrete", ds=1 / 90, chi_n=15.0)
    rng = np.random.default_rng(3)
    w0 = {"A": rng.standard_normal(nx), "B": rng.standard_normal(nx)}
    calc = scft.SCFT(p)
    calc.run(initial_fields=w0)
    capsys.readouterr()

    plan = golden("plans")["diblock_f0.5_ds1_90|discrete|1"]
    inp = plan["input"]
    M = int(np.prod(nx))
    grid = O.Grid(nx, lx)
    solver = O.Solver(grid, plan, inp["ds"], inp["bonds"])
    am = O.AndersonMixing(2 * M, 20, 1e-2, 0.1, 0.1)
    chi = np.array([[0.0, 15.0], [15.0, 0.0]])
    P = np.identity(2) - np.ones((2, 2)) @ np.linalg.inv(chi) / np.sum(np.linalg.inv(chi))
    w = np.stack([w0["A"].ravel(), w0["B"].ravel()])
    V = grid.volume
    for i in range(2):
        w[i] -= grid.integral(w[i]) / V
    err = 1e20
    for it in range(8):
        solver.compute_statistics({"A": w[0], "B": w[1]})
        phi = np.stack([solver.get_total_concentration("A"), solver.get_total_concentration("B")])
        wd = chi @ phi - P @ w
        for i in range(2):
            wd[i] -= grid.integral(wd[i]) / V
        old, err = err, np.sqrt(sum(grid.inner_product(wd[i], wd[i]) for i in range(2)) /
                                (1.0 + sum(grid.inner_product(w[i], w[i]) for i in range(2))))
        w = np.reshape(am.calculate_new_fields(w.ravel(), wd.ravel(), old, err), (2, M))
        for i in range(2):
            w[i] -= grid.integral(w[i]) / V
    np.testing.assert_allclose(calc.w, w, rtol=0, atol=1e-9)
    np.testing.assert_allclose(calc.phi["A"], phi[0], rtol=0, atol=1e-9)


def c2_fields(n, L, chi_n, seed):
    from tests.test_gpu_config_shapes import cosine_fields

    return cosine_fields(n, L, chi_n, seed)


@pytest.mark.parametrize("model,plan_key", [("continuous", "diblock_f0.36_ds0.01|continuous|1"),
                                            ("discrete", "diblock_f0.36_ds0.01|discrete|1")])
def test_c2_scft_at_config_shape(capsys, model, plan_key):
    """C2: AB diblock f=0.36, chiN=20, ds=0.01 (discrete: N=100), 64^3, L=3.3,
    AM(history 20, start 1e-2, mix_min 0.01, mix_init 0.1) on [w, L] with box
    relaxation (scft.py:617-757). From the 8-cosine start, the first three iterations
    (fields and box) against the oracle with the reference's residual, stress, mixing
    and 10 % box clip restated in numpy. The 8-cosine start is no ordered phase and
    does not reach 1e-8 in thousands of iterations (its box drifts); the run to 1e-8
    starts from the double-gyroid fields of examples/scft/Gyroid.py at 64^3 instead,
    and its final fields and box meet the tolerance when the oracle evaluates them."""
    from langevin_fts_amd import scft

    n, L, chi_n = 64, 3.3, 20.0
    nx = [n] * 3
    say(f"C2 scft {model} {nx}")
    fields = c2_fields(n, L, chi_n, 7)
    p = params(nx, [L] * 3, True, 3, f=0.36, model=model, ds=0.01, chi_n=chi_n)
    p["optimizer"] = {"name": "am", "max_hist": 20, "start_error": 1e-2, "mix_min": 0.01, "mix_init": 0.1}
    calc = scft.SCFT(p)
    calc.run(initial_fields=fields)
    capsys.readouterr()
    w_dev, lx_dev = np.array(calc.w), np.array(calc.cb.get_lx())

    plan = golden("plans")[plan_key]
    inp = plan["input"]
    M = n ** 3
    S = 2
    am = O.AndersonMixing(S * M + 3, 20, 1e-2, 0.01, 0.1)
    chi = np.array([[0.0, chi_n], [chi_n, 0.0]])
    P = np.identity(2) - np.ones((2, 2)) @ np.linalg.inv(chi) / np.sum(np.linalg.inv(chi))

    def evaluate(w, lx):
        grid = O.Grid(nx, lx)
        solver = O.Solver(grid, plan, inp["ds"], inp["bonds"])
        solver.compute_statistics({"A": w[0], "B": w[1]})
        solver.compute_stress()
        phi = np.stack([solver.get_total_concentration("A"), solver.get_total_concentration("B")])
        wd = chi @ phi - P @ w
        for i in range(S):
            wd[i] -= grid.integral(wd[i]) / grid.volume
        err = np.sqrt(sum(grid.inner_product(wd[i], wd[i]) for i in range(S)) /
                      (1.0 + sum(grid.inner_product(w[i], w[i]) for i in range(S))))
        stress = np.array(solver.get_stress())
        return grid, wd, err + np.sqrt(np.sum(stress ** 2)), stress

    w = np.stack([fields["A"].ravel(), fields["B"].ravel()])
    lx = np.array([L] * 3)
    grid = O.Grid(nx, lx)
    for i in range(S):
        w[i] -= grid.integral(w[i]) / grid.volume
    err = 1e20
    for it in range(3):
        grid, wd, e, stress = evaluate(w, lx)
        old, err = err, e
        new = am.calculate_new_fields(np.concatenate([w.ravel(), lx]), np.concatenate([wd.ravel(), -stress]), old, err)
        w = np.reshape(new[:S * M], (S, M))
        lx = (1 + np.clip((new[-3:] - lx) / lx, -0.1, 0.1)) * lx
        grid = O.Grid(nx, lx)
        for i in range(S):
            w[i] -= grid.integral(w[i]) / grid.volume
    say(f"  3 iterations: |dw| {np.max(np.abs(w_dev - w)):.1e} (|w| up to {np.max(np.abs(w)):.0f}), "
        f"box {lx_dev} / oracle {lx}")
    np.testing.assert_allclose(w_dev, w, rtol=0, atol=1e-11 * np.max(np.abs(w)))
    np.testing.assert_allclose(lx_dev, lx, rtol=0, atol=1e-12)

    p["max_iter"] = 3000
    calc = scft.SCFT(p)
    w_a, w_b = gyroid_fields(nx)
    calc.run(initial_fields={"A": w_a, "B": w_b})
    out = capsys.readouterr().out
    n_iter = sum(1 for line in out.split("\n") if re.match(r"\s+\d+\s+[-+]?\d", line))
    w_fin, lx_fin = np.array(calc.w), np.array(calc.cb.get_lx())
    _, _, e_fin, stress_fin = evaluate(w_fin, lx_fin)
    say(f"  converged in {n_iter} iterations, box {lx_fin}, oracle error of the result {e_fin:.2e}, "
        f"stress {stress_fin}")
    assert n_iter < 3000
    # the driver stopped below 1e-8; the oracle's error of the same fields differs by rounding
    assert e_fin < 1e-8 * (1 + 1e-6)
