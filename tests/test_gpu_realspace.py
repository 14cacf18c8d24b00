"""Real-space Crank-Nicolson solver (continuous chains, FP64) on the GPU against
the oracle and the reference's known answers.

Bar: 1e-10 relative (max |difference| / max |oracle|) for propagators, Q and phi;
the reference's literals (11 significant digits) to 1e-10 absolute.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def lf():
    from langevin_fts_amd import langevinfts
    return langevinfts


@pytest.fixture(scope="module")
def capi():
    from langevin_fts_amd import capi as c
    return c


@pytest.fixture(scope="module")
def ctx(capi):
    return capi.Context(0)


def api_solver(lf, nx, lx, bc, ds, bonds, polymers, mask=None, platform="hip"):
    factory = lf.PlatformSelector.create_factory(platform, False)
    cb = factory.create_computation_box(nx, lx, bc=bc if bc else None, mask=mask)
    mol = factory.create_molecules_information("continuous", ds, bonds)
    for vf, blocks, qinit in polymers:
        if qinit:
            mol.add_polymer(vf, blocks, {int(k): v for k, v in qinit.items()})
        else:
            mol.add_polymer(vf, blocks)
    an = factory.create_propagator_analyzer(mol, False)
    return cb, mol, an, factory.create_realspace_solver(cb, mol, an)


def test_realspace_kat_through_api(lf):
    """TestRealLinearContinuous3D.cpp known answers, through create_realspace_solver"""
    k = golden("realspace")["linear_continuous_3d"]
    say("realspace kat 5x4x3 periodic")
    cb, mol, an, solver = api_solver(lf, k["nx"], k["lx"], k["bc"], k["ds"], k["bonds"], k["polymers"])
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    solver.compute_concentrations()
    assert abs(solver.get_total_partition(0) - k["Q"]) < 1e-12
    for cp in k["chain_propagators"]:
        n = mol.get_polymer(cp["p"]).get_block(cp["v"], cp["u"]).n_segment
        q = solver.get_chain_propagator(cp["p"], cp["v"], cp["u"], n)
        assert np.max(np.abs(q - np.array(cp["ref"]))) < 1e-10
    for m, ref in k["phi"].items():
        assert np.max(np.abs(solver.get_total_concentration(m) - np.array(ref))) < 1e-10
    assert solver.check_total_partition()


@pytest.mark.parametrize("case", range(9))
def test_crank_nicolson_cases_match_oracle(lf, case):
    """the inputs and boundary conditions of TestCrankNicolson{1,2,3}D.cpp"""
    c = golden("realspace")["crank_nicolson_inputs"][case]
    say(f"crank-nicolson {c['source']}")
    cb, mol, an, solver = api_solver(lf, c["nx"], c["lx"], c["bc"], c["ds"], c["bonds"], c["polymers"])
    w = {m: np.array(v) for m, v in c["w"].items()}
    solver.compute_propagators(w)
    solver.compute_concentrations()
    ref = O.Solver(O.Grid(c["nx"], c["lx"]), golden("plans")["linear|continuous|0"], c["ds"], c["bonds"],
                   realspace_bc=c["bc"])
    ref.compute_statistics(w)
    assert abs(solver.get_total_partition(0) / ref.get_total_partition(0) - 1) < 1e-10
    for v, u in [(1, 2), (1, 0)]:
        n = mol.get_polymer(0).get_block(v, u).n_segment
        assert rel(solver.get_chain_propagator(0, v, u, n), ref.get_chain_propagator(0, v, u, n)) < 1e-10
    for m in ("A", "B"):
        assert rel(solver.get_total_concentration(m), ref.get_total_concentration(m)) < 1e-10


CAPI_CASES = [
    # plan, grid, lx, boundary conditions, masked
    ("branched", [12, 10, 9], [3.2, 2.8, 2.6], ["absorbing", "reflecting", "periodic", "periodic",
                                                "reflecting", "reflecting"], False),
    ("star_abc_homo_a", [16, 14, 12], [3.0, 2.8, 2.5], ["reflecting", "reflecting", "absorbing", "absorbing",
                                                        "periodic", "periodic"], True),
    ("grafted", [40, 30], [4.0, 5.0], ["periodic", "periodic", "absorbing", "reflecting"], True),
    ("with_solvent", [33], [4.0], ["absorbing", "reflecting"], False),
    ("diblock_f0.36_ds0.01", [40, 36, 44], [4.0, 3.6, 4.4], ["absorbing", "absorbing", "reflecting",
                                                            "reflecting", "periodic", "periodic"], False),
    ("mixture", [9, 7, 8], [2.0, 1.8, 1.9], ["periodic"] * 6, False),
]


@pytest.mark.parametrize("name,nx,lx,bc,masked", CAPI_CASES)
def test_realspace_matches_oracle(ctx, capi, name, nx, lx, bc, masked):
    say(f"realspace {name} {nx} {bc} mask={masked}")
    plan = golden("plans")[f"{name}|continuous|0"]
    inp = plan["input"]
    M = int(np.prod(nx))
    rng = np.random.default_rng(17)
    mask = None
    if masked:
        mask = np.ones(M)
        mask[rng.choice(M, size=M // 10, replace=False)] = 0.0
    w = {m: 0.5 * rng.standard_normal(M) for m in inp["bonds"]}
    q_init = {"G": 0.5 + rng.random(M)} if name == "grafted" else None
    solvents = [tuple(s) for s in inp["solvents"]]
    ref = O.Solver(O.Grid(nx, lx, mask), plan, inp["ds"], inp["bonds"], solvents, realspace_bc=bc)
    ref.compute_statistics(w, q_init)
    g = capi.Grid(ctx, nx, lx, mask)
    P = capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"], solvents)
    s = capi.Solver(g, P, capi.FTS_FP64, realspace_bc=bc)
    s.compute_statistics(w, q_init)
    for p in range(len(plan["polymers"])):
        assert abs(s.get_total_partition(p) / ref.get_total_partition(p) - 1) < 1e-10
    for m in inp["bonds"]:
        assert rel(s.get_total_concentration(m), ref.get_total_concentration(m)) < 1e-10
    for key in [p["key"] for p in plan["propagators"]][:4]:
        n = [p["max_n_segment"] for p in plan["propagators"] if p["key"] == key][0]
        assert rel(s.propagator(key, n), ref.propagator(key, n)) < 1e-10
    for i in range(len(solvents)):
        assert abs(s.get_solvent_partition(i) / ref.get_solvent_partition(i) - 1) < 1e-10


def test_realspace_errors(lf):
    say("realspace errors")
    k = golden("realspace")["linear_continuous_3d"]
    cb, mol, an, solver = api_solver(lf, k["nx"], k["lx"], ["reflecting"] * 6, k["ds"], k["bonds"], k["polymers"])
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    with pytest.raises(RuntimeError, match="does not support stress"):
        solver.compute_stress()
    factory = lf.PlatformSelector.create_factory("hip", False)
    cb = factory.create_computation_box(k["nx"], k["lx"])
    mol = factory.create_molecules_information("discrete", k["ds"], k["bonds"])
    mol.add_polymer(1.0, k["polymers"][0][1])
    an = factory.create_propagator_analyzer(mol, False)
    with pytest.raises(RuntimeError, match="continuous"):
        factory.create_realspace_solver(cb, mol, an)
    # the pseudo-spectral solver still refuses walls
    cb = factory.create_computation_box(k["nx"], k["lx"], bc=["absorbing"] * 6)
    mol = factory.create_molecules_information("continuous", k["ds"], k["bonds"])
    mol.add_polymer(1.0, k["polymers"][0][1])
    an = factory.create_propagator_analyzer(mol, False)
    with pytest.raises(RuntimeError, match="periodic"):
        factory.create_pseudospectral_solver(cb, mol, an)


def test_realspace_fp32_platform_runs_fp64(lf):
    """on "hip-fp32" the real-space solver still runs in FP64 and meets the FP64 bar"""
    k = golden("realspace")["linear_continuous_3d"]
    say("realspace on hip-fp32")
    cb, mol, an, solver = api_solver(lf, k["nx"], k["lx"], k["bc"], k["ds"], k["bonds"], k["polymers"],
                                     platform="hip-fp32")
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    assert abs(solver.get_total_partition(0) - k["Q"]) < 1e-12


def test_realspace_grid_beyond_fft_sizes(lf):
    """a grid the FFT plans cannot transform (prime 37 > 31) still runs in real space"""
    say("realspace 37x8")
    nx, lx, bc = [37, 8], [3.7, 0.8], ["absorbing", "absorbing", "periodic", "periodic"]
    rng = np.random.default_rng(3)
    w = {"A": 0.3 * rng.standard_normal(37 * 8), "B": 0.3 * rng.standard_normal(37 * 8)}
    blocks = [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]]
    cb, mol, an, solver = api_solver(lf, nx, lx, bc, 0.25, {"A": 1.0, "B": 1.0}, [[1.0, blocks, {}]])
    solver.compute_propagators(w)
    solver.compute_concentrations()
    ref = O.Solver(O.Grid(nx, lx), golden("plans")["linear|continuous|0"], 0.25, {"A": 1.0, "B": 1.0},
                   realspace_bc=bc)
    ref.compute_statistics(w)
    assert abs(solver.get_total_partition(0) / ref.get_total_partition(0) - 1) < 1e-10
    assert rel(solver.get_total_concentration("A"), ref.get_total_concentration("A")) < 1e-10


def test_plain_and_tiled_sweeps_agree():
    """the one-thread-per-line sweeps (lines above 1024 points, forced here by
    FTS_REAL_PLAIN_SWEEPS) and the tiled sweeps give the same propagators"""
    import os

    from langevin_fts_amd import capi

    say("realspace plain vs tiled sweeps")
    plan = golden("plans")["grafted|continuous|0"]
    inp = plan["input"]
    nx, lx = [20, 18, 23], [2.0, 1.8, 2.3]
    bc = ["absorbing", "reflecting", "periodic", "periodic", "reflecting", "absorbing"]
    M = int(np.prod(nx))
    rng = np.random.default_rng(4)
    w = {"A": 0.5 * rng.standard_normal(M)}
    q_init = {"G": 0.5 + rng.random(M)}
    ctx = capi.Context(0)
    out = {}
    for plain in (False, True):
        if plain:
            os.environ["FTS_REAL_PLAIN_SWEEPS"] = "1"
        try:
            g = capi.Grid(ctx, nx, lx)
            P = capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"])
            s = capi.Solver(g, P, capi.FTS_FP64, realspace_bc=bc)
        finally:
            os.environ.pop("FTS_REAL_PLAIN_SWEEPS", None)
        s.compute_statistics(w, q_init)
        out[plain] = (s.get_total_partition(0), s.propagator("{G}A", 100), s.get_total_concentration("A"))
    assert abs(out[True][0] / out[False][0] - 1) < 1e-12
    assert rel(out[True][1], out[False][1]) < 1e-12
    assert rel(out[True][2], out[False][2]) < 1e-12
