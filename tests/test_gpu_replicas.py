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
