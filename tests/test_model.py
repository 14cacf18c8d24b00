"""The host model layer reproduces the reference's propagator keys, analyzer
output and schedules exactly (dumps of the reference's own Polymer,
PropagatorAnalyzer and Scheduler, compiled by oracle/Makefile.ref)."""
import pytest

from tests.conftest import golden

lf = pytest.importorskip("langevin_fts_amd.langevinfts")


def build(entry, model, agg):
    inp = entry["input"]
    mol = lf.Molecules(model, inp["ds"], inp["bonds"])
    for vf, blocks, qinit in inp["polymers"]:
        if qinit:
            mol.add_polymer(vf, blocks, {int(k): v for k, v in qinit.items()})
        else:
            mol.add_polymer(vf, blocks)
    for vf, m in inp.get("solvents", []):
        mol.add_solvent(vf, m)
    factory = lf.PlatformSelector.create_factory("hip")
    return mol, factory.create_propagator_analyzer(mol, bool(agg))


PLAN_NAMES = sorted(golden("plans"))


@pytest.mark.parametrize("name", PLAN_NAMES)
def test_analyzer_and_scheduler_match_reference(name):
    entry = golden("plans")[name]
    _, model, agg = name.split("|")
    mol, an = build(entry, model, int(agg))
    assert mol.get_model_name() == entry["model"]
    for p, pref in enumerate(entry["polymers"]):
        pc = mol.get_polymer(p)
        assert pc.get_alpha() == pref["alpha"]
        assert pc.get_n_segment_total() == pref["n_segment_total"]
        for b, bref in zip(pc.get_blocks(), pref["blocks"]):
            assert (b.monomer_type, b.n_segment, b.v, b.u) == (bref["monomer"], bref["n_segment"], bref["v"], bref["u"])
            assert pc.get_propagator_key(b.v, b.u) == bref["key_vu"]
            assert pc.get_propagator_key(b.u, b.v) == bref["key_uv"]
    props = an.get_computation_propagators()
    assert list(props.keys()) == [p["key"] for p in entry["propagators"]]
    for key, ref in zip(props, entry["propagators"]):
        e = props[key]
        assert e.max_n_segment == ref["max_n_segment"]
        assert e.monomer_type == ref["monomer"]
        assert e.height == ref["height"]
        assert [list(d) for d in e.deps] == ref["deps"]
        assert sorted(e.junction_ends) == ref["junction_ends"]
    blocks = an.get_computation_blocks()
    assert [list(k) for k in blocks.keys()] == [[b["polymer"], b["left"], b["right"]] for b in entry["blocks"]]
    for k, ref in zip(blocks, entry["blocks"]):
        b = blocks[k]
        assert (b.monomer_type, b.n_segment_left, b.n_segment_right, b.n_repeated) == \
            (ref["monomer"], ref["n_left"], ref["n_right"], ref["n_repeated"])
        assert [list(vu) for vu in b.v_u] == ref["v_u"]
    for streams, sched in entry["schedules"].items():
        got = an.get_schedule(int(streams))
        assert [[list(j) for j in sl] for sl in got] == sched, streams


def test_key_parsing_matches_reference():
    mol = lf.Molecules("continuous", 0.1, {"A": 1.0})
    mol.add_polymer(1.0, [["A", 1.0, 0, 1]])
    an = lf.PlatformSelector.create_factory("hip").create_propagator_analyzer(mol, False)
    for key, ref in golden("keys").items():
        assert an.get_monomer_type_from_key(key) == ref["monomer"]
        assert [list(d) for d in an.get_deps_from_key(key)] == ref["deps"]


def test_reference_propagator_code_counts():
    """TestPropagatorCode.cpp:99-160: 29 distinct keys for the 19-block molecule."""
    entry = golden("plans")["propagator_code|continuous|0"]
    _, an = build(entry, "continuous", 0)
    assert len(an.get_computation_propagators()) == 29


@pytest.mark.parametrize("blocks,msg", [
    ([["A", 0.5, 0, 1], ["B", 0.5, 1, 2], ["A", 0.5, 2, 0]], "cycle"),
    ([["A", 0.5, 0, 1], ["B", 0.5, 2, 3]], "no route"),
    ([["A", 0.55, 0, 1]], "not an integer"),
    ([["C", 0.5, 0, 1]], "not in bond_lengths"),
    ([["A", 0.5, 0, 1], ["A", 0.5, 0, 1]], "duplicated edges"),
    ([["A", 0.5, 1, 1]], "different integers"),
])
def test_polymer_validation_errors(blocks, msg):
    """Polymer.cpp:14-187 rejects cycles, disconnected graphs, non-integer lengths,
    unknown monomers, duplicated and self edges."""
    mol = lf.Molecules("discrete", 0.1, {"A": 1.0, "B": 1.0})
    with pytest.raises(RuntimeError, match=msg):
        mol.add_polymer(1.0, blocks)


def test_invalid_model_and_platform():
    with pytest.raises(RuntimeError, match="invalid chain model"):
        lf.Molecules("gaussian", 0.1, {"A": 1.0})
    with pytest.raises(RuntimeError, match="Could not find platform"):
        lf.PlatformSelector.create_factory("cuda")
    assert lf.PlatformSelector.avail_platforms()[0] == "hip"
