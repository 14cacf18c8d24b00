"""Write the golden fixtures under tests/golden/ (run in the build container only).

Two sources, both data:

1. Known-answer literals held by the reference's own C++ tests (input fields,
   reference propagator end-slices, concentrations, partition functions and
   stresses).  The source files are read as text and only the numeric literals
   and block tables are kept; the scalar expectations are transcribed with the
   file:line they come from.
2. Dumps from oracle/_ref/ref_dump, the reference's own common-layer sources
   compiled with oracle/ref_harness.cpp: PropagatorCode keys, PropagatorAnalyzer
   propagators/blocks, Scheduler schedules, Pseudo Boltzmann/Fourier tables,
   SimpsonRule coefficients, key parsing and an Anderson-mixing trajectory.

Usage:  make -f oracle/Makefile.ref && python tests/golden/make_fixtures.py
"""
import json
import os
import re
import subprocess
import sys

REF = os.environ.get("FTS_REFERENCE_TREE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DUMP = os.path.join(REPO, "oracle", "_ref", "ref_dump")


def _read(name):
    with open(os.path.join(REF, "tests", name)) as f:
        return f.read()


def _strip_comments(src):
    return re.sub(r"//[^\n]*", "", src)


def arrays(src):
    """All `double NAME[M] = {...};` literal tables with more than one entry."""
    out = {}
    for m in re.finditer(r"double\s+(\w+)\[M\]\s*=\s*\{([^}]*)\}", _strip_comments(src)):
        vals = [float(x) for x in m.group(2).replace("\n", " ").split(",") if x.strip()]
        if len(vals) > 1:
            out[m.group(1)] = vals
    return out


def block_tables(src, env):
    """All `std::vector<BlockInput> NAME = {...};` tables, lengths evaluated in env."""
    out = {}
    src = _strip_comments(src)
    for m in re.finditer(r"std::vector<BlockInput>\s+(\w+)\s*=\s*\{(.*?)\};", src, re.S):
        blocks = []
        for b in re.finditer(r"\{\s*\"(\w+)\"\s*,\s*([^,{}]+?)\s*,\s*(\d+)\s*,\s*(\d+)\s*\}", m.group(2)):
            length = float(eval(b.group(2), {}, dict(env)))
            blocks.append([b.group(1), length, int(b.group(3)), int(b.group(4))])
        out[m.group(1)] = blocks
    return out


def run_dump(spec):
    p = subprocess.run([DUMP], input=spec, capture_output=True, text=True, check=True)
    return json.loads(p.stdout)


def molecule_spec(model, ds, bonds, polymers, solvents=(), aggregate=0, streams=(1, 2, 4, 8)):
    lines = ["mode plan", f"model {model}", f"ds {ds!r}"]
    for k, v in bonds.items():
        lines.append(f"bond {k} {v!r}")
    for vf, blocks, qinit in polymers:
        lines.append(f"polymer {vf!r}")
        for b in blocks:
            lines.append(f"block {b[0]} {b[1]!r} {b[2]} {b[3]}")
        for node, name in (qinit or {}).items():
            lines.append(f"qinit {node} {name}")
    for vf, m in solvents:
        lines.append(f"solvent {vf!r} {m}")
    lines.append(f"aggregate {aggregate}")
    lines.append("streams " + " ".join(str(s) for s in streams))
    return "\n".join(lines) + "\n"


GRID_KAT = {"nx": [5, 4, 3], "lx": [4.0, 3.0, 2.0]}


def kat_fixtures():
    kats = {}
    # TestPseudoLinear{Continuous,Discrete}3D.cpp: diblock f=0.5, N=4 (ds=1/4).
    for model, fname, qv, stress, q_lines in [
        ("continuous", "TestPseudoLinearContinuous3D.cpp", 14.899629822584,
         [-0.000708638, -0.000935597, -0.00192901], "Q :247, stress :273-285"),
        ("discrete", "TestPseudoLinearDiscrete3D.cpp", 14.9276505263205,
         [-0.000473764, -0.0006085, -0.00148826], "Q :246, stress :272-284"),
    ]:
        src = _read(fname)
        a = arrays(src)
        blk = block_tables(src, {"f": 0.5})
        kats["linear_" + model] = {
            "source": f"tests/{fname} ({q_lines})",
            **GRID_KAT, "model": model, "ds": 0.25, "bonds": {"A": 1.0, "B": 1.0},
            "polymers": [[1.0, blk["blocks"], {}]], "solvents": [],
            "w": {"A": a["w_a"], "B": a["w_b"]},
            "QV": [qv], "stress": stress,
            "chain_propagators": [
                {"p": 0, "v": 1, "u": 2, "n": "block", "ref": a["q1_last_ref"]},
                {"p": 0, "v": 1, "u": 0, "n": "block", "ref": a["q2_last_ref"]}],
            "phi": {"A": a["phi_a_ref"], "B": a["phi_b_ref"]},
            "tol_abs": 1e-7,
        }
    # TestPseudoBranched{Continuous,Discrete}3D.cpp: 19-block branched, ds=0.15, b_B=1.5.
    for model, fname, qv, line in [
        ("continuous", "TestPseudoBranchedContinuous3D.cpp", 1.5701353236e-03, 280),
        ("discrete", "TestPseudoBranchedDiscrete3D.cpp", 1.5943694728e-03, 282),
    ]:
        src = _read(fname)
        a = arrays(src)
        blk = block_tables(src, {})
        kats["branched_" + model] = {
            "source": f"tests/{fname} (Q :{line})",
            **GRID_KAT, "model": model, "ds": 0.15, "bonds": {"A": 1.0, "B": 1.5},
            "polymers": [[1.0, blk["blocks"], {}]], "solvents": [],
            "w": {"A": a["w_a"], "B": a["w_b"]},
            "QV": [qv], "stress": None,
            "chain_propagators": [
                {"p": 0, "v": 1, "u": 4, "n": "block", "ref": a["q_1_4_last_ref"]},
                {"p": 0, "v": 1, "u": 0, "n": "block", "ref": a["q_1_0_last_ref"]}],
            "phi": {"A": a["phi_a_ref"], "B": a["phi_b_ref"]},
            "tol_abs": 1e-7,
        }
    # TestPseudoBranchedMixture{Continuous,Discrete}3D.cpp: 3 polymers A/B/C.
    for model, fname, qvs in [
        ("continuous", "TestPseudoBranchedMixtureContinuous3D.cpp",
         [1.1763668568e+05, 1.4384268278e+03, 1.3821533787e+03]),
        ("discrete", "TestPseudoBranchedMixtureDiscrete3D.cpp",
         [1.3999194661e+05, 1.5625863384e+03, 1.6167175694e+03]),
    ]:
        src = _read(fname)
        a = arrays(src)
        blk = block_tables(src, {})
        kats["mixture_" + model] = {
            "source": f"tests/{fname} (Q :382-394, relative)",
            **GRID_KAT, "model": model, "ds": 0.15, "bonds": {"A": 1.0, "B": 2.0, "C": 1.5},
            "polymers": [[0.5, blk["block_input_1"], {}], [0.3, blk["block_input_2"], {}],
                         [0.2, blk["block_input_3"], {}]],
            "solvents": [],
            "w": {"A": a["w_a"], "B": a["w_b"], "C": a["w_c"]},
            "QV": qvs, "QV_relative": True, "stress": None, "chain_propagators": [],
            "phi": {"A": a["phi_a_ref"], "B": a["phi_b_ref"], "C": a["phi_c_ref"]},
            "tol_abs": 1e-7,
        }
    # TestPseudoSolvent{Continuous,Discrete}3D.cpp: a 1-segment A homopolymer equals an A solvent.
    for model, fname in [("continuous", "TestPseudoSolventContinuous3D.cpp"),
                         ("discrete", "TestPseudoSolventDiscrete3D.cpp")]:
        src = _read(fname)
        a = arrays(src)
        blk = block_tables(src, {"f": 0.5})
        kats["solvent_" + model] = {
            "source": f"tests/{fname} (:144-188)",
            **GRID_KAT, "model": model, "ds": 0.01, "bonds": {"A": 1.0, "B": 1.0},
            "w": {"A": a["w_a"], "B": a["w_b"]},
            "polymers_homopolymer": [[0.6, blk["blocks_1"], {}], [0.4, blk["blocks_2"], {}]],
            "polymers_solvent": [[0.6, blk["blocks_1"], {}]],
            "solvent": [0.4, "A"],
            "tol_Q": 1e-7, "tol_phi": 1e-4, "tol_stress": 1e-7,
        }
    return kats


def _ints(src, names):
    return [int(re.search(r"const int %s\{(\d+)\}" % n, src).group(1)) for n in names]


def _bc_lists(src):
    """`std::vector<std::string> bc_xxx = {...};` boundary-condition tables."""
    out = {}
    for m in re.finditer(r"std::vector<std::string>\s+(bc_\w+)\s*=\s*\{(.*?)\};", _strip_comments(src), re.S):
        out[m.group(1)] = re.findall(r"\"(\w+)\"", m.group(2))
    return out


def realspace_fixtures():
    """Known answers of the reference's real-space tests, the inputs of its
    Crank-Nicolson boundary-condition tests, and FiniteDifference tables."""
    out = {}
    src = _read("TestCpuTridiagonal.cpp")
    a = arrays(src)
    out["tridiagonal"] = {"source": "tests/TestCpuTridiagonal.cpp (:10-38)", "l": a["a"], "d": a["b"], "h": a["c"],
                          "rhs": a["y"], "x": a["x_answer"], "x_periodic": a["x_periodic_answer"], "tol_abs": 1e-7}
    src = _read("TestRealLinearContinuous3D.cpp")
    a = arrays(src)
    blk = block_tables(src, {"f": 0.5})
    out["linear_continuous_3d"] = {
        "source": "tests/TestRealLinearContinuous3D.cpp (q :228-244, Q :247-248, phi :254-268)",
        "nx": _ints(src, ["II", "JJ", "KK"]), "lx": [4.0, 3.0, 2.0], "bc": [], "ds": 0.25,
        "bonds": {"A": 1.0, "B": 1.0}, "polymers": [[1.0, blk["blocks"], {}]],
        "w": {"A": a["w_a"], "B": a["w_b"]}, "Q": 0.621996847395786,
        "chain_propagators": [
            {"p": 0, "v": 1, "u": 2, "n": "block", "ref": a["q1_last_ref"]},
            {"p": 0, "v": 1, "u": 0, "n": "block", "ref": a["q2_last_ref"]}],
        "phi": {"A": a["phi_a_ref"], "B": a["phi_b_ref"]}, "tol_abs": 1e-7}
    cases = []
    for fname, names, lx in [("TestCrankNicolson1D.cpp", ["II"], [4.0]),
                             ("TestCrankNicolson2D.cpp", ["II", "JJ"], None),
                             ("TestCrankNicolson3D.cpp", ["II", "JJ", "KK"], [4.0, 3.0, 2.0])]:
        src = _read(fname)
        a = arrays(src)
        blk = block_tables(src, {"f": 0.5})
        if lx is None:
            lx = [float(re.search(r"L%s\s*=\s*([0-9.]+);" % c, src).group(1)) for c in "xy"]
        for name, bc in _bc_lists(src).items():
            cases.append({"source": f"tests/{fname} ({name})", "nx": _ints(src, names), "lx": lx, "bc": bc,
                          "ds": 0.25, "bonds": {"A": 1.0, "B": 1.0}, "polymers": [[1.0, blk["blocks"], {}]],
                          "w": {"A": a["w_a"], "B": a["w_b"]}})
    out["crank_nicolson_inputs"] = cases
    lap = []
    for nx, lx, bc in [([5, 4, 3], [4.0, 3.0, 2.0], ["reflecting", "absorbing", "periodic", "periodic",
                                                     "absorbing", "reflecting"]),
                       ([7, 6], [3.0, 2.5], ["absorbing", "absorbing", "reflecting", "reflecting"]),
                       ([9], [2.0], ["reflecting", "reflecting"]),
                       ([6, 5, 4], [1.5, 2.5, 3.5], [])]:
        spec = "mode laplacian\nds 0.15\nbond A 1.0\nbond B 1.5\n" + "nx " + " ".join(map(str, nx)) + \
            "\nlx " + " ".join(map(repr, lx)) + "\n" + ("bc " + " ".join(bc) + "\n" if bc else "")
        lap.append({"nx": nx, "lx": lx, "bc": bc, "ds": 0.15, **run_dump(spec)})
    out["laplacian"] = lap
    return out


def molecule_set():
    """Molecules whose analyzer/scheduler output is pinned: every test molecule
    with a block table plus the benchmark molecules."""
    mols = {}
    kats = kat_fixtures()
    for name in ["linear_continuous", "branched_continuous", "mixture_continuous"]:
        k = kats[name]
        mols[name.rsplit("_", 1)[0]] = {"ds": k["ds"], "bonds": k["bonds"], "polymers": k["polymers"]}
    s = kats["solvent_continuous"]
    mols["solvent_pair"] = {"ds": s["ds"], "bonds": s["bonds"], "polymers": s["polymers_homopolymer"]}
    mols["with_solvent"] = {"ds": s["ds"], "bonds": s["bonds"], "polymers": s["polymers_solvent"],
                            "solvents": [s["solvent"]]}
    src = _read("TestPropagatorCode.cpp")
    mols["propagator_code"] = {"ds": 0.1, "bonds": {"A": 1.0, "B": 2.0},
                               "polymers": [[1.0, block_tables(src, {})["blocks"], {}]]}
    src = _read("TestAggregation3D.cpp")
    mols["aggregation3d"] = {"ds": 0.1, "bonds": {"A": 1.0, "B": 1.0},
                             "polymers": [[1.0, block_tables(src, {"f": 0.2})["blocks"], {}]]}
    src = _read("TestAggregationSideChains.cpp")
    bt = block_tables(src, {})
    mols["side_chains"] = {"ds": 0.1, "bonds": {"A": 1.0, "B": 0.7},
                           "polymers": [[0.2, bt["blocks_1"], {}], [0.2, bt["blocks_2"], {}],
                                        [0.2, bt["blocks_3"], {}], [0.2, bt["blocks_4"], {}]]}
    src = _read("TestAggregationDendritic.cpp")
    bt = block_tables(src, {})
    mols["dendritic"] = {"ds": 0.1, "bonds": {"A": 1.0, "B": 0.7},
                         "polymers": [[0.5, bt["blocks_1"], {}], [0.3, bt["blocks_2"], {}],
                                      [0.2, bt["blocks_3"], {}]]}
    mols["grafted"] = {"ds": 0.01, "bonds": {"A": 1.0},
                       "polymers": [[1.0, [["A", 1.0, 0, 1]], {0: "G"}]]}
    # Benchmark molecules: diblocks and the ABC star + A homopolymer mixture.
    mols["diblock_f0.5_ds0.01"] = {"ds": 0.01, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]], {}]]}
    mols["diblock_f0.36_ds0.01"] = {"ds": 0.01, "bonds": {"A": 1.0, "B": 1.0},
                                    "polymers": [[1.0, [["A", 0.36, 0, 1], ["B", 0.64, 1, 2]], {}]]}
    mols["diblock_f0.5_ds1_90"] = {"ds": 1.0 / 90, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]], {}]]}
    mols["diblock_f0.4_ds0.01"] = {"ds": 0.01, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.4, 0, 1], ["B", 0.6, 1, 2]], {}]]}
    mols["diblock_f0.4_ds1_90"] = {"ds": 1.0 / 90, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.4, 0, 1], ["B", 0.6, 1, 2]], {}]]}
    mols["diblock_f0.5_ds1_16"] = {"ds": 1.0 / 16, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]], {}]]}
    mols["diblock_f0.5_ds1_64"] = {"ds": 1.0 / 64, "bonds": {"A": 1.0, "B": 1.0},
                                   "polymers": [[1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]], {}]]}
    mols["star_abc_homo_a"] = {"ds": 0.01, "bonds": {"A": 1.0, "B": 1.0, "C": 1.2},
                               "polymers": [[0.8, [["A", 0.3, 0, 1], ["B", 0.3, 0, 2], ["C", 0.4, 0, 3]], {}],
                                            [0.2, [["A", 0.5, 0, 1]], {}]]}
    return mols


def driver_records():
    """Rows the reference's example drivers record in their trailing comments:
    fts/Lamella.py (L-FTS: iterations, mass error, Q, H, saddle error for the
    initial saddle point and Langevin steps 1-5) and scft/Gyroid.py (SCFT: iteration,
    mass error, Q, H, error, box for iterations 1-5). Only the numbers are kept."""
    num = r"([-+]?\d+\.?\d*(?:[eE][-+]?\d+)?)"
    out = {}
    path = os.path.join(REF, "examples", "fts", "Lamella.py")
    with open(path) as f:
        lines = f.read().split("\n")
    rows = []
    for ln, line in enumerate(lines, 1):
        m = re.match(r"#\s+(\d+)\s+" + num + r"\s+\[\s*" + num + r"\s*\]\s+" + num + r"\s+\[\s*" + num + r"\s*\]", line)
        if m:
            rows.append({"line": ln, "iter": int(m.group(1)), "mass_error": float(m.group(2)),
                         "Q": float(m.group(3)), "H": float(m.group(4)), "error": float(m.group(5))})
    out["lamella_lfts"] = {"source": "examples/fts/Lamella.py", "rows": rows}
    path = os.path.join(REF, "examples", "scft", "Gyroid.py")
    with open(path) as f:
        lines = f.read().split("\n")
    rows = []
    for ln, line in enumerate(lines, 1):
        m = re.match(r"\s*#\s+(\d+)\s+" + num + r"\s+\[\s*" + num + r"\s*\]\s+" + num + r"\s+" + num +
                     r"\s+\[\s*" + num + r",\s*" + num + r",\s*" + num + r"\s*\]", line)
        if m:
            rows.append({"line": ln, "iter": int(m.group(1)), "mass_error": float(m.group(2)),
                         "Q": float(m.group(3)), "H": float(m.group(4)), "error": float(m.group(5)),
                         "lx": [float(m.group(6)), float(m.group(7)), float(m.group(8))]})
    out["gyroid_scft"] = {"source": "examples/scft/Gyroid.py", "rows": rows}
    return out


def main():
    if not os.path.exists(DUMP):
        sys.exit("build oracle/_ref first: make -f oracle/Makefile.ref")
    kats = kat_fixtures()
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(kats, f)

    plans = {}
    for name, mol in molecule_set().items():
        for model in ["continuous", "discrete"]:
            for agg in [0, 1]:
                spec = molecule_spec(model, mol["ds"], mol["bonds"], mol["polymers"],
                                     mol.get("solvents", ()), agg)
                d = run_dump(spec)
                d["input"] = {"ds": mol["ds"], "bonds": mol["bonds"], "polymers": mol["polymers"],
                              "solvents": mol.get("solvents", [])}
                plans[f"{name}|{model}|{agg}"] = d
    with open(os.path.join(HERE, "plans.json"), "w") as f:
        json.dump(plans, f)

    # Pseudo tables on 3D, 2D and 1D grids (1D/2D map to leading unit axes).
    pseudo = {}
    for nx, lx, ds in [([5, 4, 3], [4.0, 3.0, 2.0], 0.25), ([12, 9, 10], [3.2, 3.3, 3.4], 0.01),
                       ([8, 6], [3.0, 2.5], 0.1), ([9], [2.0], 0.1), ([6, 5, 7], [1.5, 2.5, 3.5], 0.15)]:
        spec = "mode pseudo\n" + f"ds {ds!r}\nbond A 1.0\nbond B 1.5\n" + \
            "nx " + " ".join(map(str, nx)) + "\nlx " + " ".join(map(repr, lx)) + "\n"
        pseudo["x".join(map(str, nx))] = {"nx": nx, "lx": lx, "ds": ds, **run_dump(spec)}
    with open(os.path.join(HERE, "pseudo.json"), "w") as f:
        json.dump(pseudo, f)

    ns = list(range(0, 13)) + [35, 44, 45, 50, 63, 64, 89, 99, 100]
    simpson = run_dump("mode simpson\nsimpson " + " ".join(map(str, ns)) + "\n")
    with open(os.path.join(HERE, "simpson.json"), "w") as f:
        json.dump(simpson, f)

    keys = set()
    for d in plans.values():
        for p in d.get("propagators", []):
            keys.add(p["key"])
    keys = sorted(keys)
    deps = run_dump("mode deps\n" + "".join(f"key {k}\n" for k in keys))
    with open(os.path.join(HERE, "keys.json"), "w") as f:
        json.dump(deps, f)

    am = {}
    for name, (n_var, hist, start, mmin, minit, n_iter) in {
            "hist20": (96, 20, 1e-1, 0.1, 0.1, 40),
            "hist5_late": (64, 5, 1e-3, 0.01, 0.1, 40)}.items():
        am[name] = {"params": [n_var, hist, start, mmin, minit, n_iter],
                    **run_dump(f"mode am\nam {n_var} {hist} {start!r} {mmin!r} {minit!r} {n_iter}\n")}
    with open(os.path.join(HERE, "am.json"), "w") as f:
        json.dump(am, f)
    with open(os.path.join(HERE, "realspace.json"), "w") as f:
        json.dump(realspace_fixtures(), f)
    drivers = driver_records()
    with open(os.path.join(HERE, "drivers.json"), "w") as f:
        json.dump(drivers, f, indent=1)
    print("wrote", len(kats), "KATs,", len(plans), "plans,", len(pseudo), "pseudo grids,",
          len(keys), "keys,", len(am), "AM trajectories")


if __name__ == "__main__":
    main()
