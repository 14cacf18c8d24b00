"""Reduce-memory form (create_factory(name, True)): propagator slices and the
Anderson-mixing history in pinned host memory, read and written in place by the
same kernels. Results must equal the device-resident form exactly, and the
reference's known answers hold as they do for its own reduce-memory variants
(TestPseudoLinearContinuous3D.cpp:194-199 runs CudaComputationReduceMemory*)."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def lf():
    from langevin_fts_amd import langevinfts
    return langevinfts


def build(lf, platform, reduce, nx, lx, model, ds, bonds, polymers, aggregate=False, realspace_bc=None):
    factory = lf.PlatformSelector.create_factory(platform, reduce)
    cb = factory.create_computation_box(nx, lx, bc=realspace_bc)
    mol = factory.create_molecules_information(model, ds, bonds)
    for vf, blocks, qinit in polymers:
        mol.add_polymer(vf, blocks)
    an = factory.create_propagator_analyzer(mol, aggregate)
    if realspace_bc:
        solver = factory.create_realspace_solver(cb, mol, an)
    else:
        solver = factory.create_pseudospectral_solver(cb, mol, an)
    return factory, cb, mol, an, solver


@pytest.mark.parametrize("name", ["linear_continuous", "linear_discrete", "branched_continuous",
                                  "branched_discrete", "mixture_discrete"])
def test_reduce_memory_kats(lf, kats, name):
    say(f"reduce-memory kat {name}")
    k = kats[name]
    _, cb, mol, an, solver = build(lf, "hip", True, k["nx"], k["lx"], k["model"], k["ds"], k["bonds"],
                                   k["polymers"])
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    solver.compute_concentrations()
    V = float(np.prod(k["lx"]))
    for p, qv in enumerate(k["QV"]):
        Q = solver.get_total_partition(p)
        assert abs(Q - qv / V) / (abs(Q) if k.get("QV_relative") else 1.0) < 1e-10
    for m, ref in k["phi"].items():
        assert np.max(np.abs(solver.get_total_concentration(m) - np.array(ref))) < 1e-10
    for cp in k["chain_propagators"]:
        n = mol.get_polymer(cp["p"]).get_block(cp["v"], cp["u"]).n_segment
        assert np.max(np.abs(solver.get_chain_propagator(cp["p"], cp["v"], cp["u"], n) - np.array(cp["ref"]))) < 1e-10
    if k["stress"]:
        solver.compute_stress()
        assert np.max(np.abs(np.array(solver.get_stress()) - np.array(k["stress"]))) < 1e-8


@pytest.mark.parametrize("platform,model", [("hip", "discrete"), ("hip", "continuous"), ("hip-fp32", "discrete")])
def test_reduce_memory_equals_device_form(lf, platform, model):
    """same kernels, storage elsewhere: every output identical; the propagators leave device memory"""
    say(f"reduce-memory vs device {platform} {model}")
    nx, lx = [32, 24, 20], [4.0, 3.2, 2.8]
    blocks = [["A", 0.3, 0, 1], ["B", 0.4, 1, 2], ["A", 0.3, 1, 3], ["B", 0.2, 3, 4]]
    rng = np.random.default_rng(3)
    M = int(np.prod(nx))
    w = {"A": 0.7 * rng.standard_normal(M), "B": 0.7 * rng.standard_normal(M)}
    out = {}
    for reduce in (False, True):
        _, cb, mol, an, solver = build(lf, platform, reduce, nx, lx, model, 1 / 40, {"A": 1.0, "B": 1.2},
                                       [[1.0, blocks, {}]], aggregate=True)
        solver.compute_propagators(w)
        solver.compute_concentrations()
        solver.compute_stress()
        out[reduce] = {
            "Q": solver.get_total_partition(0),
            "phiA": solver.get_total_concentration("A"), "phiB": solver.get_total_concentration("B"),
            "stress": np.array(solver.get_stress()),
            "q": solver.get_chain_propagator(0, 1, 2, mol.get_polymer(0).get_block(1, 2).n_segment),
            "mem": solver.get_memory_usage(),
        }
    a, b = out[False], out[True]
    assert a["Q"] == b["Q"]
    for key in ("phiA", "phiB", "stress", "q"):
        assert np.array_equal(a[key], b[key]), key
    dev0, host0 = a["mem"]
    dev1, host1 = b["mem"]
    say(f"  device form: {dev0 / 2**20:.1f} MiB device, {host0 / 2**20:.1f} MiB host; "
        f"reduce-memory: {dev1 / 2**20:.1f} MiB device, {host1 / 2**20:.1f} MiB host")
    assert host0 == 0
    assert host1 > 0 and dev1 + host1 == dev0


def test_reduce_memory_realspace(lf):
    """the real-space solver's reduce-memory form: TestRealLinearContinuous3D.cpp literals"""
    k = golden("realspace")["linear_continuous_3d"]
    say("reduce-memory real space")
    _, cb, mol, an, solver = build(lf, "hip", True, k["nx"], k["lx"], "continuous", k["ds"], k["bonds"],
                                   k["polymers"], realspace_bc=["periodic"] * 6)
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    solver.compute_concentrations()
    assert abs(solver.get_total_partition(0) - k["Q"]) < 1e-12
    for m, ref in k["phi"].items():
        assert np.max(np.abs(solver.get_total_concentration(m) - np.array(ref))) < 1e-10
    assert solver.get_memory_usage()[1] > 0


def test_reduce_memory_anderson_mixing(lf):
    """CudaAndersonMixingReduceMemory: the same trajectory as the device-resident history"""
    say("reduce-memory anderson mixing")
    n_var = 5000
    rng = np.random.default_rng(9)
    A = rng.uniform(0.2, 0.9, n_var)
    target = rng.standard_normal(n_var)
    traj = {}
    for reduce in (False, True):
        am = lf.PlatformSelector.create_factory("hip", reduce).create_anderson_mixing(n_var, 8, 1.0, 0.05, 0.1)
        wv = np.zeros(n_var)
        old = 1e10
        xs = []
        for _ in range(30):
            d = A * (target - wv)
            err = float(np.sqrt(np.dot(d, d) / n_var))
            wv = am.calculate_new_fields(wv, d, old, err)
            old = err
            xs.append(wv.copy())
        traj[reduce] = np.array(xs)
    assert np.array_equal(traj[False], traj[True])
