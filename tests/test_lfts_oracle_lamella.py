"""The AB L-FTS oracle (oracle/lfts_oracle.py over the C oracle) against the rows
the reference's own example run recorded (examples/fts/Lamella.py:94-108): 32^3,
continuous AB diblock f=0.5, ds=1/16, chiN=20, L=8.0, nbar=1024, dt=8, tolerance
1e-4, Anderson mixing (20, 0.5, 0.01, 0.01); initial fields N(0, 1) from
np.random.seed(12345), Langevin noise from PCG64(12345). The initial saddle point
and Langevin steps 1-5 must give the recorded iteration counts and, to the printed
digits, Q, H and the saddle-point error. This pins the oracle's saddle point and
Leimkuhler-Matthews step directly, not only through the product. CPU only."""
import numpy as np

from oracle import oracle as O
from oracle.lfts_oracle import ABLangevinOracle
from tests.conftest import golden


def test_oracle_reproduces_lamella_rows():
    rec = golden("drivers")["lamella_lfts"]["rows"]
    plan = golden("plans")["diblock_f0.5_ds1_16|continuous|1"]
    inp = plan["input"]
    nx, lx = [32, 32, 32], [8.0, 8.0, 8.0]
    M, V = 32 ** 3, 512.0
    chi_n, dt, nbar = 20.0, 8.0, 1024
    np.random.seed(12345)
    w_A = np.random.normal(0.0, 1.0, nx).ravel()
    w_B = np.random.normal(0.0, 1.0, nx).ravel()
    w_aux = np.stack([(w_A - w_B) / 2, (w_A + w_B) / 2])
    oracle = ABLangevinOracle(O.Solver(O.Grid(nx, lx), plan, inp["ds"], inp["bonds"]),
                              O.AndersonMixing(M, 20, 5e-1, 0.01, 0.01), chi_n, [(1.0, 1.0)], 1e-4, 100)
    rng = np.random.Generator(np.random.PCG64(12345))
    sigma = np.sqrt(2 * dt * M / (V * np.sqrt(nbar)))   # lfts.py:17-18
    noise_prev = np.zeros(M)
    rows = []
    for step in range(len(rec)):
        if step > 0:
            noise = rng.normal(0.0, sigma, [1, M])[0]
            oracle.langevin_step(w_aux, phi, noise_prev, noise, dt)
            noise_prev = noise
        phi, H, n_iter, err = oracle.find_saddle_point(w_aux)
        rows.append((n_iter, oracle.solver.get_total_partition(0), H, err))
    for (n_iter, Q, H, err), ref in zip(rows, rec):
        print(f"  {n_iter} {Q:.7e} {H:.9f} {err:.7e}  recorded {ref['iter']} {ref['Q']:.7e} {ref['H']:.9f} "
              f"{ref['error']:.7e}")
        assert n_iter == ref["iter"]
        assert abs(Q / ref["Q"] - 1) < 1e-6
        assert abs(H - ref["H"]) < 2e-8
        assert abs(err / ref["error"] - 1) < 1e-5
