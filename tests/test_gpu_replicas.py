"""Replica-batched L-FTS: n independent runs advanced together (their propagators
in the same launches, one Anderson mixing each) reproduce n separate runs with the
same seeds and initial fields."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def params(nx, lx, tmp):
    return {
        "nx": nx, "lx": lx, "chain_model": "discrete", "ds": 1 / 20,
        "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": 14.0},
        "distinct_polymers": [{"volume_fraction": 1.0,
                               "blocks": [{"type": "A", "length": 0.5}, {"type": "B", "length": 0.5}]}],
        "langevin": {"max_step": 0, "dt": 0.5, "nbar": 1e3},
        "recording": {"dir": str(tmp), "recording_period": 10 ** 9, "sf_computing_period": 10 ** 9,
                      "sf_recording_period": 10 ** 9},
        "saddle": {"max_iter": 200, "tolerance": 1e-6},
        "am": {"max_hist": 20, "start_error": 1e-1, "mix_min": 0.1, "mix_init": 0.1},
        "verbose_level": 0,
    }


def test_replicas_match_separate_runs(tmp_path):
    from langevin_fts_amd import lfts

    nx, lx = [16, 12, 10], [3.2, 2.8, 2.4]
    n_rep, seed, n_steps = 3, 40, 3
    say(f"replicas {n_rep} x {nx}")
    M = int(np.prod(nx))
    rng = np.random.default_rng(8)
    init = rng.standard_normal((n_rep, 2, M))

    rep = lfts.LFTS(params(nx, lx, tmp_path), random_seed=seed, n_replicas=n_rep)
    rep.saddle_device.set_fields(np.ascontiguousarray([rep.mpt.to_aux_fields(init[r]) for r in range(n_rep)]))
    rep.saddle_device.set_noise_prev(np.zeros((n_rep, 1, M)))
    its0, _, _, _ = rep.saddle_device.find_saddle_point(1e-6, 200, 0)
    rep_its = [its0]
    for _ in range(n_steps):
        its, errs, ok = rep.replica_step()
        assert all(ok)
        rep_its.append(its)
    rep_w = rep.saddle_device.get_fields()

    for r in range(n_rep):
        one = lfts.LFTS(params(nx, lx, tmp_path), random_seed=seed + r)
        dev = one.saddle_device
        dev.set_fields(np.ascontiguousarray(one.mpt.to_aux_fields(init[r])))
        dev.set_noise_prev(np.zeros((1, M)))
        n0, _, _, _ = dev.find_saddle_point(1e-6, 200, 0)
        its = [n0]
        for _ in range(n_steps):
            noise = one.random.normal(0.0, one.langevin["sigma"], [1, M])
            dev.backup()
            dev.langevin_update(noise, one.langevin["dt"])
            n, err, _, _ = dev.find_saddle_point(1e-6, 200, 0)
            assert err < 1e-6
            its.append(n)
        say(f"  replica {r}: iterations {[x[r] for x in rep_its]} / separate {its}")
        assert [x[r] for x in rep_its] == its
        np.testing.assert_allclose(rep_w[r], dev.get_fields(), rtol=0, atol=1e-10)


def test_c4b_replicas_at_config_shape(tmp_path):
    """C4b: 32 replicas of the discrete N=64, f=0.5, chiN=17 diblock at 48^3, L=4.0,
    nbar 1e4, dt 0.8, tolerance 1e-4 with the reference's L-FTS mixing: two batched
    Langevin steps equal 32 separate runs (replica r seeded seed + r), and replica
    0's saddle points and Langevin step equal the CPU oracle's."""
    from langevin_fts_amd import lfts
    from oracle import oracle as O
    from oracle.lfts_oracle import ABLangevinOracle
    from tests.conftest import golden

    nx, lx = [48, 48, 48], [4.0, 4.0, 4.0]
    n_rep, seed, n_steps, tol = 32, 100, 2, 1e-4
    say(f"C4b replicas {n_rep} x {nx}")
    M = int(np.prod(nx))

    def c4b_params():
        p = params(nx, lx, tmp_path)
        p["ds"] = 1 / 64
        p["chi_n"] = {"A,B": 17.0}
        p["langevin"] = {"max_step": 0, "dt": 0.8, "nbar": 1e4}
        p["saddle"] = {"max_iter": 200, "tolerance": tol}
        p["am"] = {"max_hist": 20, "start_error": 5e-1, "mix_min": 0.01, "mix_init": 0.01}
        return p

    rng = np.random.default_rng(48)
    init = rng.standard_normal((n_rep, 2, M))
    rep = lfts.LFTS(c4b_params(), random_seed=seed, n_replicas=n_rep)
    rep.saddle_device.set_fields(np.ascontiguousarray([rep.mpt.to_aux_fields(init[r]) for r in range(n_rep)]))
    rep.saddle_device.set_noise_prev(np.zeros((n_rep, 1, M)))
    its0, errs0, _, _ = rep.saddle_device.find_saddle_point(tol, 200, 0)
    assert all(e < tol for e in errs0)
    rep_its, rep_errs, rep_fields = [its0], [errs0], [rep.saddle_device.get_fields()]
    for _ in range(n_steps):
        its, errs, ok = rep.replica_step()
        assert all(ok)
        rep_its.append(its)
        rep_errs.append(errs)
        rep_fields.append(rep.saddle_device.get_fields())

    for r in range(n_rep):
        one = lfts.LFTS(c4b_params(), random_seed=seed + r)
        dev = one.saddle_device
        dev.set_fields(np.ascontiguousarray(one.mpt.to_aux_fields(init[r])))
        dev.set_noise_prev(np.zeros((1, M)))
        n0, _, _, _ = dev.find_saddle_point(tol, 200, 0)
        its = [n0]
        for _ in range(n_steps):
            noise = one.random.normal(0.0, one.langevin["sigma"], [1, M])
            dev.backup()
            dev.langevin_update(noise, one.langevin["dt"])
            n, err, _, _ = dev.find_saddle_point(tol, 200, 0)
            assert err < tol
            its.append(n)
        assert [x[r] for x in rep_its] == its, (r, [x[r] for x in rep_its], its)
        np.testing.assert_allclose(rep_fields[-1][r], dev.get_fields(), rtol=0, atol=1e-10)
    say(f"  32 replicas equal separate runs; replica 0 iterations {[x[0] for x in rep_its]}")

    # replica 0 against the oracle: initial saddle point, then one Langevin step with its noise
    plan = golden("plans")["diblock_f0.5_ds1_64|discrete|1"]
    inp = plan["input"]
    oracle = ABLangevinOracle(O.Solver(O.Grid(nx, lx), plan, inp["ds"], inp["bonds"]),
                              O.AndersonMixing(M, 20, 5e-1, 0.01, 0.01), 17.0, [(1.0, 1.0)], tol, 200)
    # the saddle points stop at a 1e-4 error: rounding differences of the evaluations
    # (1e-14) grow through the Anderson steps and the Langevin step's phi, so the
    # fields agree to about 1e-7 and the iteration counts and error levels exactly
    w_ref = np.ascontiguousarray(rep.mpt.to_aux_fields(init[0]))
    phi_ref, _, n_ref, e_ref = oracle.find_saddle_point(w_ref)
    assert n_ref == rep_its[0][0] and abs(e_ref / errs0[0] - 1) < 1e-6
    np.testing.assert_allclose(w_ref, rep_fields[0][0], rtol=0, atol=1e-8)
    g = np.random.Generator(np.random.PCG64(seed))
    noise = g.normal(0.0, rep.langevin["sigma"], [1, M])[0]
    oracle.langevin_step(w_ref, phi_ref, np.zeros(M), noise, 0.8)
    _, _, n_ref, e_ref = oracle.find_saddle_point(w_ref)
    say(f"  replica 0 vs oracle: iterations {rep_its[1][0]} / {n_ref}, "
        f"max |dw| {np.max(np.abs(w_ref - rep_fields[1][0])):.1e}")
    assert n_ref == rep_its[1][0] and abs(e_ref / rep_errs[1][0] - 1) < 1e-5
    np.testing.assert_allclose(w_ref, rep_fields[1][0], rtol=0, atol=1e-6)


@pytest.mark.parametrize("mixed", [False, True])
def test_replica_graphs_equal_launches_one_by_one(tmp_path, mixed):
    """The replica search replayed from captured HIP graphs (the mixing decisions made by
    a planning kernel on the device) against the same launches issued one by one, bit
    for bit: iteration counts, error levels, Hamiltonians and fields over a first saddle
    point, three Langevin steps (each a graph of the update and two graphs per
    iteration, replayed many times with the host's own transfers and eager work in
    between) and a search stopped unconverged at max_iter. With mixed precision the FP32
    solver's graph and the switch to FP64 take part."""
    from langevin_fts_amd import lfts

    nx, lx = [16, 12, 10], [3.2, 2.8, 2.4]
    n_rep, seed, n_steps = 4, 60, 3
    M = int(np.prod(nx))
    init = np.random.default_rng(9).standard_normal((n_rep, 2, M))
    runs = {}
    for graphs in (True, False):
        say(f"replica graphs {graphs} mixed {mixed}")
        p = params(nx, lx, tmp_path)
        p["saddle"]["mixed_precision"] = mixed
        sim = lfts.LFTS(p, random_seed=seed, n_replicas=n_rep)
        dev = sim.saddle_device
        dev.set_graphs(graphs)
        dev.set_fields(np.ascontiguousarray([sim.mpt.to_aux_fields(init[r]) for r in range(n_rep)]))
        dev.set_noise_prev(np.zeros((n_rep, 1, M)))
        its, errs, hams, _ = dev.find_saddle_point(1e-6, 200, 0)
        rec = [(list(its), list(errs), list(hams), dev.get_fields())]
        for _ in range(n_steps):
            its, errs, ok = sim.replica_step()
            assert all(ok)
            dev.field_moments()   # eager work of the program's own between the replays
            rec.append((list(its), list(errs), None, dev.get_fields()))
        # a search that stops unconverged: the Hamiltonian of the last evaluation, then
        # the last mixing step still moves the fields (lfts.py:806-867)
        dev.backup()
        dev.langevin_update(np.random.default_rng(10).normal(0.0, 1.0, (n_rep, 1, M)), 0.5)
        its, errs, hams, _ = dev.find_saddle_point(1e-12, 5, 0)
        assert list(its) == [5] * n_rep
        rec.append((list(its), list(errs), list(hams), dev.get_fields()))
        runs[graphs] = rec
    for k, (a, b) in enumerate(zip(runs[True], runs[False])):
        say(f"  search {k}: iterations {a[0]} / {b[0]}")
        assert a[0] == b[0]
        assert a[1] == b[1]
        assert a[2] == b[2] or (a[2] is None and b[2] is None)
        assert np.array_equal(a[3], b[3])
