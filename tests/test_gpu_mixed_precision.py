"""Mixed-precision saddle point (saddle.mixed_precision): iterations on an FP32
solver until the error falls below the switch value, then FP64 iterations, which
alone may declare convergence. Every returned saddle point must meet the
tolerance when its fields are evaluated again by an independent FP64 solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def params(tmp, mixed, nx=(32, 32, 32), lx=(4.0, 4.0, 4.0)):
    return {
        "nx": list(nx), "lx": list(lx), "chain_model": "discrete", "ds": 1 / 90,
        "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": 17.0},
        "distinct_polymers": [{"volume_fraction": 1.0,
                               "blocks": [{"type": "A", "length": 0.5}, {"type": "B", "length": 0.5}]}],
        "langevin": {"max_step": 0, "dt": 0.8, "nbar": 1e4},
        "recording": {"dir": str(tmp), "recording_period": 10 ** 9, "sf_computing_period": 10 ** 9,
                      "sf_recording_period": 10 ** 9},
        "saddle": {"max_iter": 200, "tolerance": 1e-4, "mixed_precision": mixed},
        "am": {"max_hist": 20, "start_error": 5e-1, "mix_min": 0.01, "mix_init": 0.01},
        "verbose_level": 0,
    }


def fp64_error(lf, nx, lx, w_aux, mpt):
    """std of the incompressibility residual of an independent FP64 evaluation"""
    factory = lf.PlatformSelector.create_factory("hip", False)
    cb = factory.create_computation_box(list(nx), list(lx))
    mol = factory.create_molecules_information("discrete", 1 / 90, {"A": 1.0, "B": 1.0})
    mol.add_polymer(1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]])
    an = factory.create_propagator_analyzer(mol, True)
    solver = factory.create_pseudospectral_solver(cb, mol, an)
    w = mpt.to_monomer_fields(w_aux)
    solver.compute_statistics({"A": w[0], "B": w[1]})
    r = solver.get_total_concentration("A") + solver.get_total_concentration("B") - 1.0
    return float(np.std(r))


def test_mixed_precision_saddle_points_meet_fp64_tolerance(tmp_path):
    from langevin_fts_amd import lfts
    from langevin_fts_amd import langevinfts as lf

    nx, lx = (32, 32, 32), (4.0, 4.0, 4.0)
    M = int(np.prod(nx))
    rng = np.random.default_rng(21)
    init = rng.standard_normal((2, M))
    runs = {}
    for mixed in (False, True):
        say(f"mixed precision {mixed}")
        sim = lfts.LFTS(params(tmp_path, mixed), random_seed=7)
        dev = sim.saddle_device
        dev.set_fields(np.ascontiguousarray(sim.mpt.to_aux_fields(init)))
        dev.set_noise_prev(np.zeros((1, M)))
        its, errs = [], []
        n, err, _ = sim._find_saddle_point_device()
        its.append(n)
        errs.append(err)
        for _ in range(12):
            noise = sim.random.normal(0.0, sim.langevin["sigma"], [1, M])
            dev.backup()
            dev.langevin_update(noise, sim.langevin["dt"])
            n, err, _ = sim._find_saddle_point_device()
            assert err < 1e-4
            its.append(n)
            errs.append(err)
        w_final = dev.get_fields()
        check = fp64_error(lf, nx, lx, w_final, sim.mpt)
        say(f"  iterations {its}; returned error {errs[-1]:.2e}; independent FP64 check {check:.2e}")
        assert check < 1e-4
        runs[mixed] = (np.array(its), w_final)
    plain, mixed = runs[False][0], runs[True][0]
    # the FP32 phase does not slow the convergence down
    assert mixed[1:].mean() <= plain[1:].mean() + 2.0
    # the two trajectories stay within the saddle-point tolerance of each other
    diff = np.max(np.abs(runs[False][1] - runs[True][1]))
    say(f"  max |w_aux(mixed) - w_aux(plain)| after 12 steps: {diff:.2e}")
    assert diff < 0.05


def test_no_fp32_convergence_on_the_last_iteration(tmp_path):
    """With max_iter ending where the FP32 phase first gets below the tolerance, the
    last iteration still evaluates in FP64: a saddle point reported as converged is
    one whose fields meet the tolerance in FP64."""
    from langevin_fts_amd import lfts
    from langevin_fts_amd import langevinfts as lf

    nx, lx = (32, 32, 32), (4.0, 4.0, 4.0)
    M = int(np.prod(nx))
    init = np.random.default_rng(21).standard_normal((2, M))
    sim = lfts.LFTS(params(tmp_path, True), random_seed=7)
    dev = sim.saddle_device
    w0 = np.ascontiguousarray(sim.mpt.to_aux_fields(init))
    dev.set_noise_prev(np.zeros((1, M)))
    dev.set_fields(w0)
    n_full, err_full, _ = sim._find_saddle_point_device()
    say(f"last-iteration switch: full run {n_full} iterations, error {err_full:.2e}")
    assert err_full < 1e-4
    for k in range(max(1, n_full - 3), n_full + 1):
        dev.set_fields(w0)
        sim.saddle["max_iter"] = k
        n, err, _ = sim._find_saddle_point_device()
        check = fp64_error(lf, nx, lx, dev.get_fields(), sim.mpt)
        say(f"  max_iter {k}: {n} iterations, reported error {err:.2e}, independent FP64 error {check:.2e}")
        if err < 1e-4:
            assert check < 1e-4
