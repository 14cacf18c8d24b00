"""The reference's own test programs, restated against the hip platform's Python
module (PlatformSelector -> factory -> solver), as a user of the reference calls it."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def lf():
    from langevin_fts_amd import langevinfts
    return langevinfts


def make(lf, k, polymers, agg, platform="hip", solvents=()):
    factory = lf.PlatformSelector.create_factory(platform, False)
    cb = factory.create_computation_box(k["nx"], k["lx"])
    mol = factory.create_molecules_information(k["model"], k["ds"], k["bonds"])
    for vf, blocks, qinit in polymers:
        mol.add_polymer(vf, blocks)
    for vf, m in solvents:
        mol.add_solvent(vf, m)
    an = factory.create_propagator_analyzer(mol, agg)
    solver = factory.create_pseudospectral_solver(cb, mol, an)
    return factory, cb, mol, an, solver


@pytest.mark.parametrize("name", ["linear_continuous", "linear_discrete", "branched_continuous",
                                  "branched_discrete", "mixture_continuous", "mixture_discrete"])
@pytest.mark.parametrize("agg", [False, True])
def test_pseudo_kats(lf, kats, name, agg):
    """TestPseudo{Linear,Branched,BranchedMixture}{Continuous,Discrete}3D.cpp"""
    say(f"api kat {name} {agg}")
    k = kats[name]
    _, cb, mol, an, solver = make(lf, k, k["polymers"], agg)
    solver.compute_propagators({m: np.array(v) for m, v in k["w"].items()})
    solver.compute_concentrations()
    V = float(np.prod(k["lx"]))
    for p, qv in enumerate(k["QV"]):
        Q = solver.get_total_partition(p)
        err = abs(Q - qv / V) / (abs(Q) if k.get("QV_relative") else 1.0)
        assert err < 1e-10
    for m, ref in k["phi"].items():
        assert np.max(np.abs(solver.get_total_concentration(m) - np.array(ref))) < 1e-10
    if not agg:
        for cp in k["chain_propagators"]:
            n = mol.get_polymer(cp["p"]).get_block(cp["v"], cp["u"]).n_segment
            q = solver.get_chain_propagator(cp["p"], cp["v"], cp["u"], n)
            assert np.max(np.abs(q - np.array(cp["ref"]))) < 1e-10
    assert solver.check_total_partition()
    if k["stress"]:
        solver.compute_stress()
        assert np.max(np.abs(np.array(solver.get_stress()) - np.array(k["stress"]))) < 1e-8


@pytest.mark.parametrize("model", ["continuous", "discrete"])
def test_solvent_equals_one_segment_homopolymer(lf, kats, model):
    """TestPseudoSolvent{Continuous,Discrete}3D.cpp:144-188"""
    say(f"api solvent {model}")
    k = kats["solvent_" + model]
    w = {m: np.array(v) for m, v in k["w"].items()}
    _, _, _, _, s1 = make(lf, k, k["polymers_homopolymer"], False)
    _, _, _, _, s2 = make(lf, k, k["polymers_solvent"], False, solvents=[k["solvent"]])
    s1.compute_statistics(w)
    s2.compute_statistics(w)
    V = float(np.prod(k["lx"]))
    assert abs(s1.get_total_partition(1) - s2.get_solvent_partition(0)) / V < 1e-7
    assert np.max(np.abs(s1.get_total_concentration(1, "A") - s2.get_solvent_concentration(0))) < 1e-4
    s1.compute_stress()
    s2.compute_stress()
    assert np.max(np.abs(np.array(s1.get_stress()) - np.array(s2.get_stress()))) < 1e-7


def test_block_concentrations_and_gce(lf, kats):
    say("api blocks/gce")
    k = kats["branched_continuous"]
    _, cb, mol, an, solver = make(lf, k, k["polymers"], False)
    solver.compute_statistics({m: np.array(v) for m, v in k["w"].items()})
    blocks = solver.get_block_concentration(0)
    assert blocks.shape == (mol.get_polymer(0).get_n_blocks(), cb.get_n_grid())
    pc = mol.get_polymer(0)
    phi_a = sum(blocks[b] for b, blk in enumerate(pc.get_blocks()) if blk.monomer_type == "A")
    assert np.max(np.abs(phi_a - solver.get_total_concentration("A"))) < 1e-13
    # GCE: phi_gce = fugacity / (vf / alpha / Q) * phi  (CpuComputationContinuous.cpp:546-577)
    Q = solver.get_total_partition(0)
    fug = 0.7
    gce = solver.get_total_concentration_gce(fug, 0, "A")
    assert np.max(np.abs(gce - fug * pc.get_alpha() * Q / pc.get_volume_fraction() * phi_a)) < 1e-12
    solver.compute_stress()
    st = np.array(solver.get_stress())
    st_gce = np.array(solver.get_stress_gce([fug]))
    assert np.max(np.abs(st_gce - st * fug * pc.get_alpha() * Q / pc.get_volume_fraction())) < 1e-12


def test_box_relaxation_path(lf):
    """set_lx + update_laplacian_operator reproduce a solver built on the new box."""
    say("api set_lx")
    entry = golden("plans")["diblock_f0.36_ds0.01|continuous|0"]
    inp = entry["input"]
    nx = [16, 12, 10]
    rng = np.random.default_rng(2)
    w = {m: rng.standard_normal(int(np.prod(nx))) for m in inp["bonds"]}
    res = []
    for lx_first in ([3.0, 2.7, 2.2], [3.3, 2.8, 2.0]):
        factory = lf.PlatformSelector.create_factory("hip")
        cb = factory.create_computation_box(nx, lx_first)
        mol = factory.create_molecules_information("continuous", inp["ds"], inp["bonds"])
        mol.add_polymer(1.0, inp["polymers"][0][1])
        solver = factory.create_pseudospectral_solver(cb, mol, factory.create_propagator_analyzer(mol, True))
        if lx_first[0] == 3.0:
            cb.set_lx([3.3, 2.8, 2.0])
            solver.update_laplacian_operator()
        solver.compute_statistics(w)
        solver.compute_stress()
        res.append((solver.get_total_partition(0), np.array(solver.get_stress()), cb.get_volume()))
    assert abs(res[0][0] / res[1][0] - 1) < 1e-13
    assert np.max(np.abs(res[0][1] - res[1][1])) < 1e-12 * np.max(np.abs(res[1][1]))
    assert abs(res[0][2] - res[1][2]) < 1e-12


def test_fp32_platform(lf, kats):
    say("api hip-fp32")
    k = kats["branched_discrete"]
    _, _, _, _, solver = make(lf, k, k["polymers"], True, platform="hip-fp32")
    solver.compute_statistics({m: np.array(v) for m, v in k["w"].items()})
    V = float(np.prod(k["lx"]))
    assert abs(solver.get_total_partition(0) * V / k["QV"][0] - 1) < 1e-5


def test_anderson_mixing_api(lf):
    say("api am")
    t = golden("am")["hist20"]
    n_var, hist, start, mmin, minit, _ = t["params"]
    am = lf.PlatformSelector.create_factory("hip").create_anderson_mixing(n_var, hist, start, mmin, minit)
    w = np.array(t["w0"])
    err = 1e20
    for step in t["steps"][:25]:
        i = np.arange(w.size)
        d = 0.5 * np.tanh(w) + 0.3 * np.cos(0.1 * i) + 0.2 * np.roll(w, -1) - w
        old, err = err, np.sqrt(np.sum(d * d) / (1.0 + np.sum(w * w)))
        wn = am.calculate_new_fields(w, d, old, err)
        np.testing.assert_allclose(wn, step["w_new"], rtol=0, atol=1e-10)
        w = np.array(step["w_new"])
