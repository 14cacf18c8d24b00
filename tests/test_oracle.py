"""Pins the CPU oracle before anything is checked against it.

Every check compares the oracle with something it did not compute itself:
the reference's known-answer literals, dumps of the reference's own compiled
code, numpy's FFT, or an invariant of the mathematics (Q(n) constant along the
contour, solvent = one-segment homopolymer, stress = d(-sum phi/alpha ln Q)/dL).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

KAT_NAMES = ["linear_continuous", "linear_discrete", "branched_continuous", "branched_discrete",
             "mixture_continuous", "mixture_discrete"]


def _solver(name, k, plans, agg):
    plan = plans[f"{name.rsplit('_', 1)[0]}|{k['model']}|{agg}"]
    g = O.Grid(k["nx"], k["lx"])
    return O.Solver(g, plan, k["ds"], k["bonds"]), plan


@pytest.mark.parametrize("name", KAT_NAMES)
@pytest.mark.parametrize("agg", [0, 1])
def test_oracle_matches_reference_kats(kats, plans, name, agg):
    """TestPseudo{Linear,Branched,BranchedMixture}{Continuous,Discrete}3D.cpp known answers."""
    k = kats[name]
    s, plan = _solver(name, k, plans, agg)
    s.compute_statistics(k["w"])
    V = float(np.prod(k["lx"]))
    for p, qv in enumerate(k["QV"]):
        Q = s.get_total_partition(p)
        err = abs(Q - qv / V) / (abs(Q) if k.get("QV_relative") else 1.0)
        # literals carry 11 significant digits
        assert err < 1e-10, (p, Q, qv / V)
    for m, ref in k["phi"].items():
        assert np.max(np.abs(s.get_total_concentration(m) - np.array(ref))) < 1e-10
    if agg == 0:
        for cp in k["chain_propagators"]:
            n = s.block_n_segment(cp["p"], cp["v"], cp["u"])
            q = s.get_chain_propagator(cp["p"], cp["v"], cp["u"], n)
            assert np.max(np.abs(q - np.array(cp["ref"]))) < 1e-10
    if k["stress"]:
        s.compute_stress()
        # stress literals carry 6 significant digits (reference tolerance 1e-7 absolute)
        assert np.max(np.abs(s.get_stress() - np.array(k["stress"]))) < 1e-8


@pytest.mark.parametrize("shape", [(5, 4, 3), (7, 11, 13), (23, 27, 25), (16, 12, 10), (9, 10), (17,), (1, 6, 8)])
def test_oracle_fft_matches_numpy(shape):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape)
    g = O.Grid(list(shape), [1.0] * len(shape))
    ref = np.fft.rfftn(a).reshape(-1)
    got = g.rfft(a)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    back = g.irfft(ref)
    assert np.max(np.abs(back - a.reshape(-1))) < 1e-13


def test_oracle_tables_match_reference_pseudo():
    """Pseudo::get_boltz_bond / get_weighted_fourier_basis from the compiled reference."""
    for name, t in golden("pseudo").items():
        g = O.Grid(t["nx"], t["lx"])
        for mon, seg in [("A", 1.0), ("B", 1.5)]:
            bb = g.boltz_bond(seg * seg, t["ds"])
            bbh = g.boltz_bond(seg * seg / 2, t["ds"])
            np.testing.assert_allclose(bb, t["monomers"][mon]["boltz_bond"], rtol=1e-15, atol=0)
            np.testing.assert_allclose(bbh, t["monomers"][mon]["boltz_bond_half"], rtol=1e-15, atol=0)
        bx, by, bz = g.fourier_basis()
        np.testing.assert_allclose(bx, t["basis_x"], rtol=1e-15, atol=0)
        np.testing.assert_allclose(by, t["basis_y"], rtol=1e-15, atol=0)
        np.testing.assert_allclose(bz, t["basis_z"], rtol=1e-15, atol=0)
        assert abs(g.volume - t["volume"]) <= 1e-15 * t["volume"]


def test_oracle_simpson_matches_reference():
    for n, coeff in golden("simpson").items():
        np.testing.assert_array_equal(O.simpson(int(n)), np.array(coeff))


def am_map(w):
    """Synthetic fixed-point map of oracle/ref_harness.cpp (dump_am)."""
    i = np.arange(w.size)
    f = 0.5 * np.tanh(w) + 0.3 * np.cos(0.1 * i) + 0.2 * np.roll(w, -1)
    d = f - w
    return d, np.sqrt(np.sum(d * d) / (1.0 + np.sum(w * w)))


@pytest.mark.parametrize("name", ["hist20", "hist5_late"])
def test_oracle_anderson_matches_reference(name):
    t = golden("am")[name]
    n_var, hist, start, mmin, minit, n_iter = t["params"]
    am = O.AndersonMixing(n_var, hist, start, mmin, minit)
    w = np.array(t["w0"])
    err = 1e20
    for step in t["steps"]:
        d, e = am_map(w)
        old, err = err, e
        assert abs(err - step["err"]) <= 1e-12 * max(1.0, step["err"])
        wn = am.calculate_new_fields(w, d, old, err)
        np.testing.assert_allclose(wn, step["w_new"], rtol=0, atol=1e-11)
        w = np.array(step["w_new"])


def _random_case(model, name="branched", agg=0, seed=3, nx=(6, 5, 4), lx=(2.5, 2.0, 1.8)):
    plans = golden("plans")
    plan = plans[f"{name}|{model}|{agg}"]
    bonds = plan["input"]["bonds"]
    rng = np.random.default_rng(seed)
    g = O.Grid(list(nx), list(lx))
    s = O.Solver(g, plan, plan["input"]["ds"], bonds, plan["input"].get("solvents", []))
    w = {m: 0.4 * rng.standard_normal(int(np.prod(nx))) for m in bonds}
    return g, s, plan, w


@pytest.mark.parametrize("model", ["continuous", "discrete"])
@pytest.mark.parametrize("agg", [0, 1])
def test_oracle_partition_constant_along_contour(model, agg):
    """check_total_partition (CpuComputationContinuous.cpp:755-817): recomputed
    here in numpy from the oracle's propagator slices, independent of its Q code."""
    g, s, plan, w = _random_case(model, "side_chains", agg)
    s.compute_statistics(w)
    V = g.volume
    dv = V / g.M
    for p in range(len(plan["polymers"])):
        vals = []
        for blk in plan["blocks"]:
            if blk["polymer"] != p:
                continue
            NL, NR, rep, nvu = blk["n_left"], blk["n_right"], blk["n_repeated"], len(blk["v_u"])
            if model == "continuous":
                ns = range(0, NR + 1)
                for n in ns:
                    a = s.propagator(blk["left"], NL - n)
                    b = s.propagator(blk["right"], n)
                    vals.append(np.sum(a * b) * dv * rep / V / nvu)
            else:
                e = np.exp(-np.asarray(w[blk["monomer"]]) * plan["input"]["ds"])
                for n in range(1, NR + 1):
                    a = s.propagator(blk["left"], NL - n + 1)
                    b = s.propagator(blk["right"], n)
                    vals.append(np.sum(a * b / e) * dv * rep / V / nvu)
        vals = np.array(vals)
        assert np.max(vals) - np.min(vals) < 1e-12 * np.max(np.abs(vals))
        assert abs(vals[0] - s.get_total_partition(p)) < 1e-12 * abs(vals[0])


@pytest.mark.parametrize("model", ["continuous", "discrete"])
def test_oracle_solvent_equals_one_segment_homopolymer(kats, model):
    """TestPseudoSolvent{Continuous,Discrete}3D.cpp:144-188."""
    k = kats["solvent_" + model]
    plans = golden("plans")
    g = O.Grid(k["nx"], k["lx"])
    s1 = O.Solver(g, plans[f"solvent_pair|{model}|0"], k["ds"], k["bonds"])
    s2 = O.Solver(g, plans[f"with_solvent|{model}|0"], k["ds"], k["bonds"], [k["solvent"]])
    s1.compute_statistics(k["w"])
    s2.compute_statistics(k["w"])
    V = float(np.prod(k["lx"]))
    assert abs(s1.get_total_partition(1) - s2.get_solvent_partition(0)) / V < 1e-7
    phi_h = s1.get_total_concentration("A", polymer=1)
    phi_s = s2.get_solvent_concentration(0)
    assert np.max(np.abs(phi_h - phi_s)) < k["tol_phi"]
    s1.compute_stress()
    s2.compute_stress()
    assert np.max(np.abs(s1.get_stress() - s2.get_stress())) < k["tol_stress"]


@pytest.mark.parametrize("model,name,tol", [("continuous", "diblock_f0.36_ds0.01", 1e-3),
                                            ("discrete", "branched", 1e-6)])
def test_oracle_stress_is_box_derivative(model, name, tol):
    """Stress = d/dL of -sum_p phi_p/alpha_p ln Q_p at fixed w (TestStressLinear3D.cpp:267-450
    checks the same relation by central differences with relative tolerance 1e-3; the
    continuous form is the Simpson quadrature of the derivative, so ds must be small)."""
    g, s, plan, w = _random_case(model, name, 0, nx=(8, 6, 5), lx=(3.2, 2.9, 2.4))
    s.compute_statistics(w)
    s.compute_stress()
    stress = s.get_stress()
    pols = plan["polymers"]

    def free(lx):
        g2 = O.Grid([8, 6, 5], lx)
        s2 = O.Solver(g2, plan, plan["input"]["ds"], plan["input"]["bonds"])
        s2.compute_propagators(w)
        return -sum(p["volume_fraction"] / p["alpha"] * np.log(s2.get_total_partition(i))
                    for i, p in enumerate(pols))

    lx0 = [3.2, 2.9, 2.4]
    for d in range(3):
        h = 1e-5 * lx0[d]
        lp, lm = list(lx0), list(lx0)
        lp[d] += h
        lm[d] -= h
        fd = (free(lp) - free(lm)) / (2 * h)
        assert abs(stress[d] - fd) <= tol * abs(fd), (d, stress[d], fd)
