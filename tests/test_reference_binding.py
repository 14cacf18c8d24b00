"""The hip platform behind the reference's own C++ factory: HipFactory and its
classes (langevin_fts_amd/csrc/reference_binding/) derive from the reference's
abstract AbstractFactory, PropagatorComputation, ComputationBox, AndersonMixing and
Array (AbstractFactory.h:21-62) and call libfts_hip.so. oracle/Makefile.binding
compiles them against the reference tree's headers and common sources into
oracle/_ref/hip_factory_check, a caller that uses only the reference's interfaces.

CPU: the binding compiles and links (only where the reference tree is present).
GPU: the built program runs the reference's known-answer tests on the hip platform
(TestPseudo{Linear,Branched}{Continuous,Discrete}3D.cpp literals, 1e-7 absolute as
there) and a 32^3 diblock against the FP64 oracle at 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import golden

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "oracle", "_ref", "hip_factory_check")
LIB = os.path.join(REPO, "langevin_fts_amd", "libfts_hip.so")


def say(msg):
    print(msg, flush=True)


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/common") or not os.path.exists(LIB),
                    reason="needs the reference tree and the built libfts_hip.so")
def test_binding_builds_against_reference_headers():
    subprocess.run(["make", "-s", "-f", os.path.join(REPO, "oracle", "Makefile.binding")], check=True, cwd=REPO)
    assert os.path.exists(EXE)
    deps = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libfts_hip.so" in deps


def write_case(tmp, model, ds, nx, lx, bonds, polymers, w, aggregate=1):
    spec = [f"model {model}", f"ds {ds!r}", "nx " + " ".join(map(str, nx)), "lx " + " ".join(map(repr, lx)),
            f"aggregate {aggregate}"]
    for m, b in bonds.items():
        spec.append(f"bond {m} {b!r}")
    for vf, blocks, *_ in polymers:
        spec.append(f"polymer {vf!r}")
        for b in blocks:
            spec.append(f"block {b[0]} {b[1]!r} {b[2]} {b[3]}")
    with open(os.path.join(tmp, "spec.txt"), "w") as f:
        f.write("\n".join(spec) + "\n")
    np.concatenate([np.asarray(w[m], dtype=np.float64).ravel() for m in sorted(bonds)]).tofile(
        os.path.join(tmp, "w.bin"))


def run_case(tmp, bonds, M):
    r = subprocess.run([EXE, os.path.join(tmp, "spec.txt"), os.path.join(tmp, "w.bin"), os.path.join(tmp, "phi.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    Q, stress = {}, None
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts and parts[0] == "Q":
            Q[int(parts[1])] = float(parts[2])
        elif parts and parts[0] == "stress":
            stress = np.array([float(x) for x in parts[1:]])
    phi = np.fromfile(os.path.join(tmp, "phi.bin")).reshape(len(bonds), M)
    return Q, stress, {m: phi[i] for i, m in enumerate(sorted(bonds))}


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/hip_factory_check is built only where the "
                                                    "reference tree is present")
@pytest.mark.parametrize("name", ["linear_continuous", "linear_discrete", "branched_continuous", "branched_discrete"])
def test_reference_interfaces_run_the_kats(tmp_path, name):
    say(f"reference binding KAT {name}")
    k = golden("kats")[name]
    write_case(str(tmp_path), k["model"], k["ds"], k["nx"], k["lx"], k["bonds"], k["polymers"], k["w"])
    Q, stress, phi = run_case(str(tmp_path), k["bonds"], int(np.prod(k["nx"])))
    V = float(np.prod(k["lx"]))
    assert abs(Q[0] * V - k["QV"][0]) < 1e-7
    for m, ref in k["phi"].items():
        assert np.max(np.abs(phi[m] - np.array(ref))) < 1e-7
    if k["stress"]:
        assert np.max(np.abs(stress - np.array(k["stress"]))) < 1e-7


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/hip_factory_check is built only where the "
                                                    "reference tree is present")
def test_reference_interfaces_match_oracle_at_32(tmp_path):
    from oracle import oracle as O

    say("reference binding 32^3 diblock vs oracle")
    key = "diblock_f0.5_ds0.01|continuous|1"
    plan = golden("plans")[key]
    inp = plan["input"]
    nx, lx = [32, 32, 32], [4.0, 4.0, 4.0]
    M = 32 ** 3
    rng = np.random.Generator(np.random.PCG64(3))
    w = {m: rng.standard_normal(M) for m in inp["bonds"]}
    write_case(str(tmp_path), plan["model"], inp["ds"], nx, lx, inp["bonds"], inp["polymers"], w)
    Q, stress, phi = run_case(str(tmp_path), inp["bonds"], M)
    ref = O.Solver(O.Grid(nx, lx), plan, inp["ds"], inp["bonds"])
    ref.compute_statistics(w)
    ref.compute_stress()
    assert abs(Q[0] / ref.get_total_partition(0) - 1) < 1e-10
    for m in inp["bonds"]:
        r = ref.get_total_concentration(m)
        assert np.max(np.abs(phi[m] - r)) < 1e-10 * np.max(np.abs(r))
    s = np.array(ref.get_stress())
    assert np.max(np.abs(stress - s)) < 1e-10 * np.max(np.abs(s))
