"""Field-theory tables of the L-FTS driver (examples/fts/lfts.py:20-262) against
the closed forms of AB and symmetric ABC melts; no GPU needed."""
import numpy as np
import pytest

from langevin_fts_amd.lfts import SymmetricPolymerTheory, calculate_sigma


def test_ab_closed_form(capsys):
    chi = 17.0
    t = SymmetricPolymerTheory(["A", "B"], {"A,B": chi})
    np.testing.assert_allclose(t.matrix_a, [[1, 1], [-1, 1]], atol=1e-14)
    np.testing.assert_allclose(t.matrix_a_inv, [[0.5, -0.5], [0.5, 0.5]], atol=1e-14)
    np.testing.assert_allclose(t.eigenvalues, [-chi, 0.0], atol=1e-12)
    assert t.aux_fields_real_idx == [0] and t.aux_fields_imag_idx == [1]
    assert abs(t.h_const - chi / 4) < 1e-13
    np.testing.assert_allclose(t.h_coef_mu1, [0.0], atol=1e-14)
    np.testing.assert_allclose(t.h_coef_mu2, [1 / chi], rtol=1e-14)
    # d/dchiN with vector_s held at its unperturbed value (lfts.py:84-105): the
    # reference energy keeps chi/4 and has no derivative; mu2 = 1/chi gives -1/chi^2
    assert abs(t.h_const_deriv_chin["A,B"]) < 1e-8
    np.testing.assert_allclose(t.h_coef_mu1_deriv_chin["A,B"], [0.0], atol=1e-8)
    np.testing.assert_allclose(t.h_coef_mu2_deriv_chin["A,B"], [-1 / chi ** 2], rtol=1e-6)


def test_ab_func_deriv_matches_closed_form():
    chi = 13.0
    t = SymmetricPolymerTheory(["A", "B"], {"A,B": chi})
    rng = np.random.default_rng(0)
    w_aux = rng.standard_normal((2, 50))
    phi = {"A": rng.random(50), "B": rng.random(50)}
    real, _ = t.compute_func_deriv(w_aux, phi, [0])
    imag, _ = t.compute_func_deriv(w_aux, phi, [1])
    np.testing.assert_allclose(real[0], 2 * w_aux[0] / chi + phi["A"] - phi["B"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(imag[0], -(phi["A"] + phi["B"] - 1.0), rtol=1e-14, atol=1e-14)


def test_abc_equal_chi_spectrum():
    """Equal chi: P chi P = -chi P, eigenvalues (-chi, -chi, 0), two real fields."""
    chi = 15.0
    t = SymmetricPolymerTheory(["A", "B", "C"], {"A,B": chi, "B,C": chi, "A,C": chi})
    np.testing.assert_allclose(np.sort(t.eigenvalues), [-chi, -chi, 0.0], atol=1e-12)
    assert t.aux_fields_real_idx == [0, 1] and t.aux_fields_imag_idx == [2]
    S = 3
    np.testing.assert_allclose(t.matrix_a @ t.matrix_a_inv, np.identity(S), atol=1e-13)
    np.testing.assert_allclose(t.matrix_a[:, 2], np.ones(S), atol=1e-14)
    # columns are eigenvectors of P chi P scaled by sqrt(S)
    pcp = t.matrix_p @ t._chi_matrix({"A,B": chi, "B,C": chi, "A,C": chi}) @ t.matrix_p
    for i in range(S):
        np.testing.assert_allclose(pcp @ t.matrix_a[:, i], t.eigenvalues[i] * t.matrix_a[:, i], atol=1e-12)
    assert np.all(t.matrix_a[0] >= 0)


def test_abc_general_func_deriv():
    """Asymmetric ABC: the linear coefficient tables the device saddle point takes
    reproduce dH/dw_aux written out term by term (lfts.py:219-245)."""
    chi_n = {"A,B": 12.0, "B,C": 18.0, "A,C": 25.0}
    t = SymmetricPolymerTheory(["A", "B", "C"], chi_n)
    assert t.aux_fields_imag_idx[-1] == 2
    np.testing.assert_allclose(t.matrix_a @ t.matrix_a_inv, np.identity(3), atol=1e-13)
    cw, cp, c0 = t.func_deriv_coefficients([0, 1, 2])
    rng = np.random.default_rng(1)
    w_aux = rng.standard_normal((3, 20))
    phi = {m: rng.random(20) for m in "ABC"}
    h, _ = t.compute_func_deriv(w_aux, phi, [0, 1, 2])
    for r, i in enumerate([0, 1, 2]):
        ref = np.zeros(20)
        if i != 2:
            ref += 2 * t.h_coef_mu2[i] * w_aux[i] + t.h_coef_mu1[i]
            for j, m in enumerate("ABC"):
                ref += t.matrix_a[j, i] * phi[m]
        else:
            ref += phi["A"] + phi["B"] + phi["C"] - 1.0
        if i in t.aux_fields_imag_idx:
            ref = -ref
        np.testing.assert_allclose(h[r], ref, rtol=1e-13, atol=1e-13)


def test_sigma():
    assert abs(calculate_sigma(1e4, 0.8, 64 ** 3, 6.0 ** 3) - np.sqrt(2 * 0.8 * 64 ** 3 / (216.0 * 100.0))) < 1e-12


def test_noise_stream_keeps_the_draw_order():
    """Noise drawn one step ahead equals direct draws of the same generator, and a
    rewind puts the generator back to the state the next step draws from."""
    from langevin_fts_amd.lfts import NoiseStream

    direct = np.random.Generator(np.random.PCG64(123))
    expected = [direct.normal(0.0, 0.3, [1, 1000]) for _ in range(6)]
    g = np.random.Generator(np.random.PCG64(123))
    stream = NoiseStream(g, 0.3, [1, 1000])
    got = [stream.next() for _ in range(3)]
    for a, b in zip(got, expected[:3]):
        np.testing.assert_array_equal(a, b)
    stream.rewind()   # one draw was pending: the state must be that after 3 draws
    check = np.random.Generator(np.random.PCG64(123))
    for _ in range(3):
        check.normal(0.0, 0.3, [1, 1000])
    assert g.bit_generator.state == check.bit_generator.state
    got += [stream.next() for _ in range(3)]
    for a, b in zip(got, expected):
        np.testing.assert_array_equal(a, b)


def test_guess_learning_rate_falls_on_a_cosine():
    """LFTS.train_guess passes lr at the first training step, (lr + lr_final)/2 half-way and lr_final at
    the last, on a cosine (checked with a stand-in for the device object)"""
    import types

    from langevin_fts_amd.lfts import LFTS

    seen = []

    class Device:
        def train_guess(self, guess, lr, iters, replay):
            seen.append(lr)
            return 0.0

    T = 100
    # a network that exists, its scale samples collected
    sim = types.SimpleNamespace(guess=object(), guess_trained_steps=0, guess_losses=[],
                                guess_scale_samples=[[1.0, 1.0]] * 8, saddle_device=Device(),
                                learned={"lr": 1e-3, "lr_final": 1e-4, "train_steps": T, "train_iters": 16,
                                         "replay": 64})
    for _ in range(T + 3):
        LFTS.train_guess(sim)
    assert seen[0] == pytest.approx(1e-3, rel=1e-12)
    assert seen[T // 2] == pytest.approx(0.5 * (1e-3 + 1e-4), rel=1e-12)
    assert seen[T] == pytest.approx(1e-4, rel=1e-12)
    assert seen[T + 2] == pytest.approx(1e-4, rel=1e-12)   # held after the training window
    assert all(a >= b for a, b in zip(seen, seen[1:]))
    k = 30
    assert seen[k] == pytest.approx(1e-4 + 0.5 * 9e-4 * (1 + np.cos(np.pi * k / T)), rel=1e-12)
