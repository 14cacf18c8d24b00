"""Pins the oracle's real-space Crank-Nicolson path (continuous chains).

Checks, each against something the oracle did not compute itself:
- the reference's tridiagonal known answers (TestCpuTridiagonal.cpp) and a dense solve;
- FiniteDifference::get_laplacian_matrix tables dumped from the reference's own source;
- the reference's real-space known answers (TestRealLinearContinuous3D.cpp);
- closed forms: on a product of per-axis eigenvectors (cosines for periodic and
  reflecting sides, sines for absorbing ones) one ADI step with w = 0 multiplies
  the mode by a number fixed by the per-axis eigenvalues;
- a dense-matrix statement of the ADI step (Kronecker products of the per-axis
  operators, numpy.linalg.solve) on random fields, masks and mixed boundaries.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden


def rs():
    return golden("realspace")


def test_oracle_tridiagonal_matches_reference():
    k = rs()["tridiagonal"]
    l, d, h, rhs = (np.array(k[n]) for n in ("l", "d", "h", "rhs"))
    x = O.tridiagonal(l, d, h, rhs)
    xp = O.tridiagonal(l, d, h, rhs, periodic=True)
    # literals carry 15 significant digits
    assert np.max(np.abs(x - np.array(k["x"]))) < 1e-12
    assert np.max(np.abs(xp - np.array(k["x_periodic"]))) < 1e-11
    n = len(rhs)
    A = np.diag(d) + np.diag(l[1:], -1) + np.diag(h[:-1], 1)
    np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=1e-12)
    A[0, n - 1] += l[0]
    A[n - 1, 0] += h[n - 1]
    np.testing.assert_allclose(xp, np.linalg.solve(A, rhs), rtol=1e-11)


@pytest.mark.parametrize("case", range(4))
def test_oracle_laplacian_matches_reference(case):
    c = rs()["laplacian"][case]
    g = O.Grid(c["nx"], c["lx"])
    for m, b in {"A": 1.0, "B": 1.5}.items():
        mine = O.laplacian(g, c["bc"], b * b, c["ds"])
        for a, ref in enumerate(c["monomers"][m]):
            for row in ("l", "d", "h"):
                np.testing.assert_allclose(mine[a][row], ref[row], rtol=1e-15, atol=0)


def _linear_solver(plans, k, bc):
    g = O.Grid(k["nx"], k["lx"])
    return O.Solver(g, plans["linear|continuous|0"], k["ds"], k["bonds"], realspace_bc=bc)


def test_oracle_realspace_matches_reference_kat(plans):
    """TestRealLinearContinuous3D.cpp: periodic box, diblock f=0.5, ds=1/4."""
    k = rs()["linear_continuous_3d"]
    s = _linear_solver(plans, k, k["bc"])
    s.compute_statistics(k["w"])
    assert abs(s.get_total_partition(0) - k["Q"]) < 1e-12
    for cp in k["chain_propagators"]:
        n = s.block_n_segment(cp["p"], cp["v"], cp["u"])
        q = s.get_chain_propagator(cp["p"], cp["v"], cp["u"], n)
        # literals carry 11 significant digits
        assert np.max(np.abs(q - np.array(cp["ref"]))) < 1e-10
    for m, ref in k["phi"].items():
        assert np.max(np.abs(s.get_total_concentration(m) - np.array(ref))) < 1e-10


def _mode_1d(kind, n, k):
    """eigenvector and eigenvalue of K (B = f/2 K) for one axis"""
    i = np.arange(n)
    if kind == "periodic":
        return np.cos(2 * np.pi * k * i / n), 2 - 2 * np.cos(2 * np.pi * k / n)
    if kind == "reflecting":
        return np.cos(np.pi * k * (i + 0.5) / n), 2 - 2 * np.cos(np.pi * k / n)
    return np.sin(np.pi * (k + 1) * (i + 1) / (n + 1)), 2 - 2 * np.cos(np.pi * (k + 1) / (n + 1))


@pytest.mark.parametrize("nx,lx", [([10], [4.0]), ([5, 4], [4.0, 3.0]), ([6, 5, 7], [1.5, 2.5, 3.5])])
@pytest.mark.parametrize("kinds", ["periodic", "reflecting", "absorbing", "mixed"])
def test_oracle_realspace_eigenmodes(plans, nx, lx, kinds):
    D = len(nx)
    axis_kinds = ([["periodic", "reflecting", "absorbing"][a % 3] for a in range(D)] if kinds == "mixed"
                  else [kinds] * D)
    bc = [b for kd in axis_kinds for b in (kd, kd)]
    ds, bond = 0.01, 1.2
    g = O.Grid(nx, lx)
    s = O.Solver(g, plans["grafted|continuous|0"], ds, {"A": bond}, realspace_bc=bc)
    for ks in [(1,) * D, (0,) * D, tuple(range(2, 2 + D))]:
        mode = np.ones(())
        betas = []
        for a in range(D):
            v, lam = _mode_1d(axis_kinds[a], nx[a], ks[a])
            mode = np.multiply.outer(mode, v)
            f = bond ** 2 * ds / (6 * (lx[a] / nx[a]) ** 2)
            betas.append(0.5 * f * lam)
        gain = (1 - betas[0] - 2 * sum(betas[1:])) / (1 + betas[0])
        for b in betas[1:]:
            gain = (gain + b) / (1 + b)
        s.compute_propagators({"A": np.zeros(g.M)}, {"G": mode.ravel()})
        for n in (1, 7, 100):
            q = s.propagator("{G}A", n)
            np.testing.assert_allclose(q, gain ** n * mode.ravel(), rtol=1e-11, atol=1e-13)


def _axis_op(t, periodic):
    n = len(t["d"])
    T = np.diag(t["d"]) + np.diag(t["l"][1:], -1) + np.diag(t["h"][:-1], 1)
    if periodic:
        T[0, n - 1] += t["l"][0]
        T[n - 1, 0] += t["h"][n - 1]
    return T - np.eye(n)


def _dense_steps(grid, bc, bond, ds, w, q0, mask, n_steps):
    """q_e = E q; x_0 = (I+B_0)^-1 (I - B_0 - 2 sum B_b) q_e; x_b = (I+B_b)^-1 (x_{b-1} + B_b q_e); E x mask"""
    nx = grid.nx
    D, M = len(nx), grid.M
    tabs = O.laplacian(grid, bc, bond * bond, ds)
    Bs = []
    for a in range(D):
        ops = [np.eye(n) for n in nx]
        ops[a] = _axis_op(tabs[a], bc[2 * a] == "periodic")
        full = ops[0]
        for op in ops[1:]:
            full = np.kron(full, op)
        Bs.append(full)
    I = np.eye(M)
    E = np.exp(-w * ds / 2)
    out = [q0 * mask]
    q = out[0]
    for _ in range(n_steps):
        qe = E * q
        x = np.linalg.solve(I + Bs[0], (I - Bs[0] - 2 * sum(Bs[1:], np.zeros((M, M)))) @ qe)
        for B in Bs[1:]:
            x = np.linalg.solve(I + B, x + B @ qe)
        q = E * x * mask
        out.append(q)
    return out


CASES = [
    ([9], [2.0], ["reflecting", "absorbing"]),
    ([5, 4], [4.0, 3.0], ["absorbing", "reflecting", "periodic", "periodic"]),
    ([4, 3, 5], [2.0, 1.5, 2.5], ["reflecting", "absorbing", "absorbing", "absorbing", "periodic", "periodic"]),
    ([3, 4, 2], [1.0, 1.5, 0.8], ["periodic"] * 6),
]


@pytest.mark.parametrize("nx,lx,bc", CASES)
@pytest.mark.parametrize("masked", [False, True])
def test_oracle_realspace_matches_dense_adi(plans, nx, lx, bc, masked):
    rng = np.random.default_rng(11)
    g0 = O.Grid(nx, lx)
    M = g0.M
    mask = np.ones(M)
    if masked:
        mask[rng.choice(M, size=max(1, M // 6), replace=False)] = 0.0
    g = O.Grid(nx, lx, mask if masked else None)
    ds, bond = 0.05, 1.3
    s = O.Solver(g, plans["grafted|continuous|0"], ds, {"A": bond}, realspace_bc=bc)
    w = rng.normal(0, 1, M)
    q0 = rng.uniform(0.5, 1.5, M)
    s.compute_propagators({"A": w}, {"G": q0})
    ref = _dense_steps(g0, bc, bond, ds, w, q0, mask, 5)
    for n in range(6):
        np.testing.assert_allclose(s.propagator("{G}A", n), ref[n], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("bc", [["reflecting"] * 6, ["periodic"] * 6])
def test_oracle_realspace_conserves_constant(plans, bc):
    """w = 0 with reflecting or periodic sides: q = 1 stays 1, so Q = 1"""
    g = O.Grid([5, 4, 3], [4.0, 3.0, 2.0])
    s = O.Solver(g, plans["linear|continuous|0"], 0.25, {"A": 1.0, "B": 1.0}, realspace_bc=bc)
    s.compute_statistics({"A": np.zeros(g.M), "B": np.zeros(g.M)})
    assert abs(s.get_total_partition(0) - 1.0) < 1e-13
    np.testing.assert_allclose(s.get_total_concentration("A") + s.get_total_concentration("B"), 1.0, rtol=1e-13)


def test_oracle_realspace_absorbing_loses_chains(plans):
    """absorbing sides remove chain ends: Q < 1 for w = 0"""
    g = O.Grid([6, 5, 4], [2.0, 2.0, 2.0])
    s = O.Solver(g, plans["linear|continuous|0"], 0.25, {"A": 1.0, "B": 1.0},
                 realspace_bc=["absorbing"] * 6)
    s.compute_statistics({"A": np.zeros(g.M), "B": np.zeros(g.M)})
    assert 0.0 < s.get_total_partition(0) < 0.95


def test_oracle_realspace_rejects_discrete(plans):
    g = O.Grid([5, 4, 3], [4.0, 3.0, 2.0])
    with pytest.raises(RuntimeError, match="continuous"):
        O.Solver(g, plans["linear|discrete|0"], 0.25, {"A": 1.0, "B": 1.0}, realspace_bc=["periodic"] * 6)


def test_realspace_fixture_cases_are_well_formed():
    for c in rs()["crank_nicolson_inputs"]:
        assert len(c["bc"]) == 2 * len(c["nx"])
        assert len(c["w"]["A"]) == int(np.prod(c["nx"]))
    assert {tuple(c["nx"]) for c in rs()["crank_nicolson_inputs"]} == {(10,), (5, 4), (5, 4, 3)}
    assert len(list(itertools.chain.from_iterable(c["bc"] for c in rs()["crank_nicolson_inputs"]))) == 36
