"""The L-FTS driver's recorded outputs (lfts.py:720-788): structure_function_%06d.mat
and dH_%06d.mat against a restatement with the AB closed forms. For two monomer
types, w_A = w- + w+ and w_B = -w- + w+ (eigenvalues -chiN and 0), so
mu_A(k) = -FFT(w-)/(2 chiN M), mu_B(k) = +FFT(w-)/(2 chiN M), phi_i(k) = FFT(phi_i)/M,
S_ij(k) = sum over the computing steps of mu_i(k) conj(phi_j(k)), recorded times
(sf_computing_period / sf_recording_period) V sqrt(nbar); H is the saddle point's
Hamiltonian and dH/dchiN = -<w-^2>/chiN^2: the coefficient of <w-^2> is 1/chiN, and
the reference differentiates the constant term with the pressure eigenvector held
fixed, which leaves nothing of it for AB (lfts.py:83-105, :265-273; a central finite
difference, hence the 1e-8 tolerance). The fields and concentrations the driver
used are captured as it runs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_structure_function_and_dh_recordings(tmp_path):
    from scipy.io import loadmat

    from langevin_fts_amd import lfts

    nx, lx, chi_n, nbar = [16, 12, 10], [3.2, 2.8, 2.4], 15.0, 1e3
    say("lfts recordings")
    params = {
        "nx": nx, "lx": lx, "chain_model": "discrete", "ds": 1 / 20,
        "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": chi_n},
        "distinct_polymers": [{"volume_fraction": 1.0,
                               "blocks": [{"type": "A", "length": 0.4}, {"type": "B", "length": 0.6}]}],
        "langevin": {"max_step": 4, "dt": 0.5, "nbar": nbar},
        "recording": {"dir": str(tmp_path), "recording_period": 10 ** 9, "sf_computing_period": 1,
                      "sf_recording_period": 2},
        "saddle": {"max_iter": 200, "tolerance": 1e-6},
        "am": {"max_hist": 20, "start_error": 1e-1, "mix_min": 0.1, "mix_init": 0.1},
        "verbose_level": 0,
    }
    sim = lfts.LFTS(params, random_seed=17)
    seen, hams = [], []
    terms = sim.structure_function_terms
    find = sim._find_saddle_point_device

    def spy_terms(w_aux, phi):
        seen.append((np.array(w_aux), {k: np.array(v) for k, v in phi.items()}))
        return terms(w_aux, phi)

    def spy_find():
        out = find()
        hams.append(out[2])
        return out

    sim.structure_function_terms = spy_terms
    sim._find_saddle_point_device = spy_find
    rng = np.random.default_rng(2)
    sim.run(initial_fields={"A": rng.standard_normal(nx), "B": rng.standard_normal(nx)})
    assert len(seen) == 4
    M, V = int(np.prod(nx)), float(np.prod(lx))
    for file_step, steps in [(2, seen[0:2]), (4, seen[2:4])]:
        got = loadmat(os.path.join(tmp_path, "structure_function_%06d.mat" % file_step))
        want = {"A_A": 0.0, "A_B": 0.0, "B_B": 0.0}
        for w_aux, phi in steps:
            wk = np.fft.rfftn(w_aux[0].reshape(nx))
            mu = {"A": -wk / (2 * chi_n * M), "B": wk / (2 * chi_n * M)}
            ph = {t: np.fft.rfftn(phi[t].reshape(nx)) / M for t in "AB"}
            for key in want:
                a, b = key.split("_")
                want[key] = want[key] + mu[a] * np.conj(ph[b])
        for key in want:
            w = want[key] * (1 / 2) * V * np.sqrt(nbar)
            g = got["structure_function_" + key]
            assert g.shape == w.shape
            assert np.max(np.abs(g - w)) <= 1e-12 * np.max(np.abs(w)), key
        dh = loadmat(os.path.join(tmp_path, "dH_%06d.mat" % file_step))
        idx = [file_step - 2, file_step - 1]
        np.testing.assert_allclose(dh["H_history"].ravel(), [hams[i + 1] for i in idx], rtol=1e-14, atol=0)
        want_dh = [-np.mean(seen[i][0][0] ** 2) / chi_n ** 2 for i in idx]
        np.testing.assert_allclose(dh["dH_history_A_B"].ravel(), want_dh, rtol=1e-8, atol=0)
    say("  structure functions and dH/dchiN recordings match the closed forms")
