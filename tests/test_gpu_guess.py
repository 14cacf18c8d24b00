"""What the learned guess gains: after online training on the run's own saddle
points, saddle points that start from the CNN's prediction need fewer Anderson
iterations than the reference's warm start, at the same tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


# warmup 0: training starts at the first Langevin step after the random start, where the
# samples are least typical (the network's scales come from the first 8 samples)
@pytest.mark.parametrize("inputs,warmup", [("residual", 10), ("delta", 10), ("residual", 0)])
def test_learned_guess_reduces_saddle_iterations(tmp_path, inputs, warmup):
    from langevin_fts_amd import lfts

    nx, lx = [32, 32, 32], [4.0, 4.0, 4.0]
    say(f"learned guess gain {nx} ({inputs}, {warmup} plain steps first)")
    params = {
        "nx": nx, "lx": lx, "chain_model": "discrete", "ds": 1 / 90,
        "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": 17.0},
        "distinct_polymers": [{"volume_fraction": 1.0,
                               "blocks": [{"type": "A", "length": 0.5}, {"type": "B", "length": 0.5}]}],
        "langevin": {"max_step": 0, "dt": 0.8, "nbar": 1e4},
        "recording": {"dir": str(tmp_path), "recording_period": 10 ** 9, "sf_computing_period": 10 ** 9,
                      "sf_recording_period": 10 ** 9},
        "saddle": {"max_iter": 200, "tolerance": 1e-4},
        "am": {"max_hist": 20, "start_error": 5e-1, "mix_min": 0.01, "mix_init": 0.01},
        "verbose_level": 0,
        "learned_guess": {"train_steps": 150, "lr": 1e-3, "train_iters": 4, "replay": 32, "inputs": inputs,
                          "seed": 0},
    }
    sim = lfts.LFTS(params, random_seed=5)
    M = int(np.prod(nx))
    rng = np.random.default_rng(5)
    sim.saddle_device.set_fields(np.ascontiguousarray(sim.mpt.to_aux_fields(rng.standard_normal((2, M)))))
    sim.saddle_device.set_noise_prev(np.zeros((1, M)))
    sim._find_saddle_point_device()

    def steps(n, learned):
        its, errs = [], []
        for _ in range(n):
            noise = sim.random.normal(0.0, sim.langevin["sigma"], [1, M])
            if learned:
                it, err, ok = sim.learned_step(noise)
            else:
                dev = sim.saddle_device
                dev.backup()
                dev.langevin_update(noise, sim.langevin["dt"])
                it, err, _ = sim._find_saddle_point_device()
                ok = err < sim.saddle["tolerance"]
                if not ok:
                    dev.restore()
            assert ok
            its.append(it)
            errs.append(err)
        return np.array(its), np.array(errs)

    steps(warmup, False)
    train_its, _ = steps(150, True)
    assert sim.guess_ready()
    plain_its, plain_err = steps(40, False)
    learned_its, learned_err = steps(40, True)
    say(f"  iterations per step: training {train_its.mean():.1f}, plain {plain_its.mean():.2f}, "
        f"learned {learned_its.mean():.2f} ({plain_its.mean() / learned_its.mean():.2f}x fewer)")
    assert np.all(plain_err < 1e-4) and np.all(learned_err < 1e-4)
    assert learned_its.mean() < 0.9 * plain_its.mean()
