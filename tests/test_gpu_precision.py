"""The FP32 form ("hip-fp32": propagators and FFTs in FP32; Q, phi, stress and the
Anderson Gram solve in FP64) against the FP64 path at the precision-form shapes:
continuous AB diblock f=0.4, ds=0.01, 96^3 (L=4.5) and 128^3 (L=6.0), fields
N(0, 2^2), and the discrete N=90 variant. The FP64 path itself matches the oracle
to 1e-10 (test_gpu_capi.py); the errors below are the FP32 form's, for Q, phi, stress
and the mean-field free energy per chain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run(lf, platform, nx, lx, model, ds, w):
    factory = lf.PlatformSelector.create_factory(platform, False)
    cb = factory.create_computation_box(nx, lx)
    mol = factory.create_molecules_information(model, ds, {"A": 1.0, "B": 1.0})
    mol.add_polymer(1.0, [["A", 0.4, 0, 1], ["B", 0.6, 1, 2]])
    solver = factory.create_pseudospectral_solver(cb, mol, factory.create_propagator_analyzer(mol, True))
    solver.compute_statistics(w)
    solver.compute_stress()
    return (solver.get_total_partition(0), solver.get_total_concentration("A"), solver.get_total_concentration("B"),
            np.array(solver.get_stress()))


@pytest.mark.parametrize("n,L,model,ds", [(96, 4.5, "continuous", 0.01), (128, 6.0, "continuous", 0.01),
                                          (96, 4.5, "discrete", 1 / 90), (128, 6.0, "discrete", 1 / 90)])
def test_fp32_form_errors(n, L, model, ds):
    from langevin_fts_amd import langevinfts as lf

    nx, lx = [n] * 3, [L] * 3
    say(f"fp32 form {n}^3 {model}")
    rng = np.random.Generator(np.random.PCG64(2024))
    w = {"A": rng.normal(0.0, 2.0, n ** 3), "B": rng.normal(0.0, 2.0, n ** 3)}
    q64, a64, b64, s64 = run(lf, "hip", nx, lx, model, ds, w)
    q32, a32, b32, s32 = run(lf, "hip-fp32", nx, lx, model, ds, w)
    q_err = abs(q32 / q64 - 1)
    phi_err = max(np.max(np.abs(a32 - a64)), np.max(np.abs(b32 - b64)))
    stress_err = np.max(np.abs(s32 - s64)) / np.max(np.abs(s64))

    # mean-field free energy per chain of the AB melt at chiN = 18 (the precision form's):
    # F = -ln Q + < chiN phi_A phi_B - w_A phi_A - w_B phi_B >
    def free_energy(q, a, b):
        return -np.log(q) + np.mean(18.0 * a * b - w["A"] * a - w["B"] * b)

    f64, f32 = free_energy(q64, a64, b64), free_energy(q32, a32, b32)
    f_err = abs(f32 - f64)
    say(f"  Q rel {q_err:.2e}  max|dphi| {phi_err:.2e} (phi max {max(a64.max(), b64.max()):.2f})  "
        f"stress rel {stress_err:.2e}  free energy {f64:.6f}, |dF| {f_err:.2e}")
    assert q_err < 1e-4
    assert phi_err < 1e-3 * max(a64.max(), b64.max())
    assert stress_err < 1e-3
    assert f_err < 1e-4 * max(1.0, abs(f64))
