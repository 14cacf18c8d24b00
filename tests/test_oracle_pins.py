"""Closed-form pins of the oracle's grafting (q_init) and mask paths
(ComputationBox.cpp:49-98 and :229-262, CpuComputationContinuous.cpp:148-369,
CpuComputationDiscrete.cpp:279-291), independent of the product. CPU only.

- A grafted chain in w = 0 started from q_init = c + cos(k x): the continuous
  Richardson step reduces to IFFT[B FFT[q]] (E = 1, B_h^2 = B), so the mode decays
  by exp(-b^2 ds k^2 / 6) per step and the constant stays; the discrete chain's
  q[1] = q_init E and q[n+1] = E IFFT[B FFT[q[n]]] give the factor to the power n-1.
- The mask enters through dv = dV mask (integrals) and through q *= mask after each
  step; the box volume is prod(L) at construction and sum(dv) after set_lx (the
  reference's own rule), and Q = sum dv q_L q_R / V with that volume."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden


@pytest.mark.parametrize("model", ["continuous", "discrete"])
def test_grafted_fourier_mode_decays_by_the_bond_factor(model):
    plan = golden("plans")[f"grafted|{model}|0"]
    inp = plan["input"]
    nx, lx = [12, 10, 8], [3.0, 2.5, 2.0]
    M = int(np.prod(nx))
    x = (np.arange(nx[0]) * lx[0] / nx[0])[:, None, None] * np.ones(nx)
    k = 2 * np.pi * 2 / lx[0]
    q0 = (1.5 + np.cos(k * x)).ravel()
    ref = O.Solver(O.Grid(nx, lx), plan, inp["ds"], inp["bonds"])
    ref.compute_propagators({"A": np.zeros(M)}, {"G": q0})
    decay = np.exp(-1.0 * inp["ds"] * k * k / 6.0)
    for n in [1, 7, 50, 100]:
        power = n if model == "continuous" else n - 1
        want = 1.5 + np.cos(k * x).ravel() * decay ** power
        np.testing.assert_allclose(ref.propagator("{G}A", n), want, rtol=0, atol=1e-12)


def test_mask_weighted_integrals_volume_and_partition():
    nx, lx = [12, 10, 8], [3.0, 2.5, 2.0]
    M = int(np.prod(nx))
    rng = np.random.default_rng(5)
    mask = (rng.random(M) > 0.2).astype(np.float64)
    a, b = rng.standard_normal((2, M))
    g = O.Grid(nx, lx, mask)
    dV = np.prod(lx) / M
    assert abs(g.integral(a) - np.sum(mask * dV * a)) < 1e-13
    assert abs(g.inner_product(a, b) - np.sum(mask * dV * a * b)) < 1e-13
    assert abs(g.volume - np.prod(lx)) < 1e-13                 # construction: prod(L)
    lx2 = [3.3, 2.2, 2.1]
    g.set_lx(lx2)
    dV2 = np.prod(lx2) / M
    assert abs(g.volume - np.sum(mask) * dV2) < 1e-12          # after set_lx: sum(dv)
    assert abs(g.integral(a) - np.sum(mask * dV2 * a)) < 1e-13

    # masked propagators vanish inside the mask, and Q uses the mask-weighted dv
    for model in ["continuous", "discrete"]:
        plan = golden("plans")[f"diblock_f0.5_ds1_90|{model}|0"]
        inp = plan["input"]
        og = O.Grid(nx, lx, mask)
        ref = O.Solver(og, plan, inp["ds"], inp["bonds"])
        w = {m: 0.4 * rng.standard_normal(M) for m in inp["bonds"]}
        ref.compute_statistics(w)
        blk = plan["blocks"][0]
        ql = ref.propagator(blk["left"], blk["n_left"])
        n_r = 0 if model == "continuous" else 1
        qr = ref.propagator(blk["right"], n_r)
        assert np.all(ql[mask == 0] == 0) and np.all(ref.propagator(blk["right"], 5)[mask == 0] == 0)
        if model == "discrete":
            qr = qr / np.exp(-w[blk["monomer"]] * inp["ds"])
        Q = np.sum(mask * dV * ql * qr) / np.prod(lx) / len(blk["v_u"]) * blk["n_repeated"]
        assert abs(Q / ref.get_total_partition(0) - 1) < 1e-12
