"""The CPU baseline in the reference CPU path's structure (oracle/cpu_pocketfft.py:
thread pool over a slice's propagators, one sequential pocketfft per job) computes
what the C oracle computes: Q, phi and the continuous stress to 1e-11."""
import numpy as np
import pytest

from oracle import cpu_pocketfft as P
from oracle import oracle as O
from tests.conftest import golden


@pytest.mark.parametrize("key", ["diblock_f0.36_ds0.01|continuous|1", "diblock_f0.5_ds1_90|discrete|1",
                                 "diblock_f0.4_ds1_90|discrete|1", "star_abc_homo_a|continuous|0"])
def test_pocketfft_baseline_matches_oracle(key):
    plan = golden("plans")[key]
    inp = plan["input"]
    nx, lx = [12, 10, 8], [3.3, 3.0, 2.7]
    M = int(np.prod(nx))
    rng = np.random.default_rng(4)
    w = {m: rng.standard_normal(M) for m in inp["bonds"]}
    ref = O.Solver(O.Grid(nx, lx), plan, inp["ds"], inp["bonds"], inp["solvents"])
    ref.compute_statistics(w)
    base = P.PocketSolver(nx, lx, plan, inp["ds"], inp["bonds"], threads=3)
    base.compute_propagators(w)
    base.compute_concentrations()
    for p in range(len(plan["polymers"])):
        assert abs(base.Q[p] / ref.get_total_partition(p) - 1) < 1e-11
    for m in inp["bonds"]:
        a, b = base.get_total_concentration(m), ref.get_total_concentration(m)
        assert np.max(np.abs(a - b)) < 1e-11 * np.max(np.abs(b))
    if plan["model"] == "continuous":
        ref.compute_stress()
        s, r = base.compute_stress(), np.array(ref.get_stress())
        assert np.max(np.abs(s - r)) < 1e-11 * np.max(np.abs(r))
