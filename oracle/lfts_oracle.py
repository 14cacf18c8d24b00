"""AB-melt L-FTS saddle point and Langevin step on the CPU oracle (test
infrastructure only: imported by tests/ as the checker, never by the product).

For two monomer types A < B the field theory of examples/fts/lfts.py:20-262 has a
closed form, used here instead of an eigen decomposition so that the product's
general tables are checked against independent arithmetic:

    w_A = w_- + w_+,  w_B = -w_- + w_+                      (matrix A = [[1, 1], [-1, 1]])
    dH/dw_-  = 2 w_- / chiN + phi_A - phi_B                  (real field, lfts.py:228-233)
    -dH/dw_+ = -(phi_A + phi_B - 1)                          (pressure, sign of lfts.py:243-244)
    H = -<w_+> + <w_-^2> / chiN - sum_p phi_p/alpha_p ln Q_p + chiN/4   (lfts.py:197-216)
    eigenvalues (-chiN, 0): dt scaling 1 for both fields     (lfts.py:498-501)

find_saddle_point follows lfts.py:806-867 and langevin_step lfts.py:701-713.
"""
import numpy as np


class ABLangevinOracle:
    def __init__(self, solver, am, chi_n, polymers, tolerance, max_iter):
        """solver: oracle.Solver over monomers A, B; am: oracle.AndersonMixing of n_var M;
        polymers: [(volume fraction, alpha)] in polymer order."""
        self.solver = solver
        self.am = am
        self.chi_n = chi_n
        self.polymers = polymers
        self.tolerance = tolerance
        self.max_iter = max_iter

    def concentrations(self, w_aux):
        self.solver.compute_statistics({"A": w_aux[0] + w_aux[1], "B": -w_aux[0] + w_aux[1]})
        return {"A": self.solver.get_total_concentration("A"), "B": self.solver.get_total_concentration("B")}

    def hamiltonian(self, w_aux):
        h = -np.mean(w_aux[1]) + np.mean(w_aux[0] ** 2) / self.chi_n
        for p, (vf, alpha) in enumerate(self.polymers):
            h -= vf / alpha * np.log(self.solver.get_total_partition(p))
        return h + self.chi_n / 4.0

    def find_saddle_point(self, w_aux):
        """Updates w_aux in place; returns (phi, H, iterations, error level)."""
        error_level = 1e20
        self.am.reset_count()
        hamiltonian = None
        for it in range(1, self.max_iter + 1):
            phi = self.concentrations(w_aux)
            h = -(phi["A"] + phi["B"] - 1.0)
            old_error_level, error_level = error_level, np.std(h)
            if error_level < self.tolerance or it == self.max_iter:
                hamiltonian = self.hamiltonian(w_aux)
            if error_level < self.tolerance:
                break
            w_aux[1] = self.am.calculate_new_fields(w_aux[1], -h, old_error_level, error_level)
        w_aux[1] -= np.mean(w_aux[1])
        return phi, hamiltonian, it, error_level

    def langevin_step(self, w_aux, phi, noise_prev, noise_cur, dt):
        lam = 2.0 * w_aux[0] / self.chi_n + phi["A"] - phi["B"]
        w_aux[0] += -lam * dt + 0.5 * (noise_prev + noise_cur)
