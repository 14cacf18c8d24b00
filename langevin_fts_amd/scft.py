"""Self-consistent field theory (SCFT) on the hip platform.

The driver of examples/scft/scft.py with the same parameters dictionary, the same
`run` / `save_results` / `compute_concentrations` methods, the same printed rows
(iteration, mass error, Q, free energy, error level, box) and optional box
relaxation by the stress (scft.py:617-757). Propagators, concentrations, stress and
Anderson mixing run on the GPU through the platform's solver; the residuals and
error levels are formed on the host as in the reference.
"""
import copy
import itertools
import json
import os
import re
import time

import numpy as np

from . import langevinfts as lf
from .lfts import SymmetricPolymerTheory


class Adam:
    """Adam step on the fields (scft.py:20-55): w += lr gamma^(t-1) m_hat / (sqrt(v_hat) + eps)."""

    def __init__(self, M, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8, gamma=1.0):
        self.M, self.lr, self.b1, self.b2, self.eps, self.gamma = M, lr, b1, b2, eps, gamma
        self.count = 1
        self.m = np.zeros(M)
        self.v = np.zeros(M)

    def reset_count(self):
        self.count = 1
        self.m[:] = 0.0
        self.v[:] = 0.0

    def calculate_new_fields(self, w_current, w_diff, old_error_level, error_level):
        lr = self.lr * self.gamma ** (self.count - 1)
        self.m = self.b1 * self.m + (1.0 - self.b1) * w_diff
        self.v = self.b2 * self.v + (1.0 - self.b2) * w_diff ** 2
        m_hat = self.m / (1.0 - self.b1 ** self.count)
        v_hat = self.v / (1.0 - self.b2 ** self.count)
        self.count += 1
        return w_current + lr * m_hat / (np.sqrt(v_hat) + self.eps)


class SCFT:
    def __init__(self, params):
        self.monomer_types = sorted(list(params["segment_lengths"].keys()))
        self.segment_lengths = copy.deepcopy(params["segment_lengths"])
        self.distinct_polymers = copy.deepcopy(params["distinct_polymers"])
        assert len(self.monomer_types) == len(set(self.monomer_types)), "There are duplicated monomer_types"

        platform = params.get("platform", "hip")
        factory = lf.PlatformSelector.create_factory(platform, False)
        factory.display_info()
        cb = factory.create_computation_box(params["nx"], params["lx"])

        self.chi_n = {}
        for pair_str, value in params["chi_n"].items():
            pair = re.split(",| |_|/", pair_str)
            assert pair[0] in self.segment_lengths, f"Monomer type '{pair[0]}' is not in 'segment_lengths'."
            assert pair[1] in self.segment_lengths, f"Monomer type '{pair[1]}' is not in 'segment_lengths'."
            assert pair[0] != pair[1], "Do not add self interaction parameter, " + pair_str + "."
            key = ",".join(sorted(pair[:2]))
            assert key not in self.chi_n, f"There are duplicated χN ({key}) parameters."
            self.chi_n[key] = value
        for pair in itertools.combinations(self.monomer_types, 2):
            self.chi_n.setdefault(",".join(sorted(pair)), 0.0)

        self.mpt = SymmetricPolymerTheory(self.monomer_types, self.chi_n)

        # residual projector P = I - J chi^-1 / sum(chi^-1) (scft.py:281-295)
        S = len(self.monomer_types)
        chi = np.zeros((S, S))
        for i, j in itertools.combinations(range(S), 2):
            key = self.monomer_types[i] + "," + self.monomer_types[j]
            if key in self.chi_n:
                chi[i, j] = chi[j, i] = self.chi_n[key]
        self.matrix_chi = chi
        chi_inv = np.linalg.inv(chi)
        self.matrix_p = np.identity(S) - np.ones((S, S)) @ chi_inv / np.sum(chi_inv)

        assert len(self.distinct_polymers) >= 1, "There is no polymer chain."
        assert np.isclose(sum(p["volume_fraction"] for p in self.distinct_polymers), 1.0), \
            "The sum of volume fractions must be equal to 1."
        for polymer in self.distinct_polymers:
            blocks = polymer["blocks"]
            linear = "v" not in blocks[0]
            inp = []
            for b in blocks:
                if linear:
                    assert "v" not in b and "u" not in b, \
                        "Index v should exist in all blocks, or it should not exist in all blocks for each polymer."
                    inp.append([b["type"], b["length"], len(inp), len(inp) + 1])
                else:
                    assert "v" in b and "u" in b, \
                        "Index v should exist in all blocks, or it should not exist in all blocks for each polymer."
                    inp.append([b["type"], b["length"], b["v"], b["u"]])
            polymer["blocks_input"] = inp

        self.random_fraction = {}
        for polymer in self.distinct_polymers:
            if not any("fraction" in b for b in polymer["blocks"]):
                continue
            assert len(polymer["blocks"]) == 1, "Only single block random copolymer is allowed."
            fractions = polymer["blocks"][0]["fraction"]
            assert np.isclose(sum(fractions.values()), 1.0), \
                "The sum of volume fractions of random copolymer must be equal to 1."
            name = polymer["blocks"][0]["type"]
            assert name not in self.segment_lengths, \
                f"The name of random copolymer '{name}' is already used as a type in 'segment_lengths' or other random copolymer"
            polymer["block_monomer_types"] = [name]
            self.segment_lengths[name] = np.sqrt(sum(self.segment_lengths[t] ** 2 * f for t, f in fractions.items()))
            self.random_fraction[name] = fractions

        molecules = factory.create_molecules_information(params["chain_model"], params["ds"], self.segment_lengths)
        for polymer in self.distinct_polymers:
            if "initial_conditions" in polymer:
                molecules.add_polymer(polymer["volume_fraction"], polymer["blocks_input"],
                                      polymer["initial_conditions"])
            else:
                molecules.add_polymer(polymer["volume_fraction"], polymer["blocks_input"])
        analyzer = factory.create_propagator_analyzer(molecules, params.get("aggregate_propagator_computation", True))
        solver = factory.create_pseudospectral_solver(cb, molecules, analyzer)

        self.scale_stress = params.get("scale_stress", 1)
        n_grid = int(np.prod(params["nx"]))
        n_var = S * n_grid + (len(params["lx"]) if params["box_is_altering"] else 0)
        opt = params["optimizer"]
        if opt["name"] == "am":
            self.field_optimizer = factory.create_anderson_mixing(n_var, opt["max_hist"], opt["start_error"],
                                                                  opt["mix_min"], opt["mix_init"])
        elif opt["name"] == "adam":
            self.field_optimizer = Adam(M=n_var, lr=opt["lr"], gamma=opt["gamma"])
        else:
            raise ValueError(f'Invalid optimizer name: {opt}. Choose among "am" and "adam".')
        max_iter = params.get("max_iter", 2000)
        tolerance = params.get("tolerance", 1e-8)

        print("---------- Simulation Parameters ----------")
        print("Platform :", platform)
        print("Box Dimension: %d" % (cb.get_dim()))
        print("Nx:", cb.get_nx())
        print("Lx:", cb.get_lx())
        print("dx:", cb.get_dx())
        print("Volume: %f" % (cb.get_volume()))
        print("Chain model: %s" % (params["chain_model"]))
        print("Segment lengths:\n\t", list(self.segment_lengths.items()))
        print("Conformational asymmetry (epsilon): ")
        for a, b in itertools.combinations(self.monomer_types, 2):
            print("\t%s/%s: %f" % (a, b, self.segment_lengths[a] / self.segment_lengths[b]))
        print("χN: ")
        for key in self.chi_n:
            print("\t%s: %f" % (key, self.chi_n[key]))
        print("P matrix for field residuals:\n\t", str(self.matrix_p).replace("\n", "\n\t"))
        for p in range(molecules.get_n_polymer_types()):
            print("distinct_polymers[%d]:" % (p))
            print("\tvolume fraction: %f, alpha: %f, N: %d" % (
                molecules.get_polymer(p).get_volume_fraction(), molecules.get_polymer(p).get_alpha(),
                molecules.get_polymer(p).get_n_segment_total()))
        analyzer.display_blocks()
        analyzer.display_propagators()

        self.params = params
        self.chain_model = params["chain_model"]
        self.ds = params["ds"]
        self.box_is_altering = params["box_is_altering"]
        self.max_iter = max_iter
        self.tolerance = tolerance
        self.cb = cb
        self.molecules = molecules
        self.propagator_analyzer = analyzer
        self.solver = solver

    def compute_concentrations(self, w):
        """phi of every monomer type for monomer fields w[S][M] (scft.py:513-541)."""
        S = len(self.monomer_types)
        w_input = {self.monomer_types[i]: w[i] for i in range(S)}
        for name, fractions in self.random_fraction.items():
            w_input[name] = np.zeros(self.cb.get_n_grid())
            for t, f in fractions.items():
                w_input[name] += w_input[t] * f
        t0 = time.time()
        self.solver.compute_propagators(w_input)
        self.solver.compute_concentrations()
        elapsed = {"pseudo": time.time() - t0}
        t0 = time.time()
        phi = {t: self.solver.get_total_concentration(t) for t in self.monomer_types}
        elapsed["phi"] = time.time() - t0
        for name, fractions in self.random_fraction.items():
            phi[name] = self.solver.get_total_concentration(name)
            for t, f in fractions.items():
                phi[t] += phi[name] * f
        return phi, elapsed

    def hamiltonian(self, w_aux, total_partitions):
        """Free energy per volume without the reference energy (scft.py:214-230)."""
        return self.mpt.compute_hamiltonian(self.molecules, w_aux, total_partitions) - self.mpt.h_const

    def save_results(self, path):
        """w, phi and the theory's matrices as .mat, .json or .yaml (scft.py:543-615)."""
        m_dic = {"initial_params": self.params,
                 "dim": self.cb.get_dim(), "nx": self.cb.get_nx(), "lx": self.cb.get_lx(),
                 "monomer_types": self.monomer_types, "chi_n": dict(self.chi_n), "chain_model": self.chain_model,
                 "ds": self.ds, "eigenvalues": self.mpt.eigenvalues, "matrix_a": self.mpt.matrix_a,
                 "matrix_a_inverse": self.mpt.matrix_a_inv}
        for i, name in enumerate(self.monomer_types):
            m_dic["w_" + name] = self.w[i]
        for name in self.monomer_types:
            m_dic["phi_" + name] = self.phi[name]
        for k in m_dic:
            if type(m_dic[k]).__module__ == np.__name__:
                m_dic[k] = m_dic[k].tolist()
        m_dic["initial_params"] = {k: (v.tolist() if type(v).__module__ == np.__name__ else v)
                                   for k, v in m_dic["initial_params"].items()}
        ext = os.path.splitext(path)[1]
        if ext == ".mat":
            from scipy.io import savemat

            savemat(path, m_dic, long_field_names=True, do_compression=True)
        elif ext == ".json":
            with open(path, "w") as f:
                json.dump(m_dic, f, indent=2)
        elif ext == ".yaml":
            import yaml

            with open(path, "w") as f:
                f.write(yaml.dump(m_dic, default_flow_style=None, width=90, sort_keys=False))
        else:
            raise ValueError("Invalid output file extension.")

    def run(self, initial_fields, q_init=None):
        """Field relaxation to the self-consistent solution (scft.py:617-757)."""
        S = len(self.monomer_types)
        M = self.cb.get_n_grid()
        energy_total = 1.0e20
        error_level = 1.0e20
        self.field_optimizer.reset_count()
        print("---------- Run ----------")
        print("iteration, mass error, total_partitions, energy_total, error_level", end="")
        print(", box size" if self.box_is_altering else "")

        w = np.zeros([S, M])
        for i in range(S):
            w[i] = np.reshape(initial_fields[self.monomer_types[i]], M)
        for i in range(S):
            w[i] -= self.cb.integral(w[i]) / self.cb.get_volume()

        for scft_iter in range(1, self.max_iter + 1):
            phi, _ = self.compute_concentrations(w)
            w_aux = self.mpt.to_aux_fields(w)
            total_partitions = [self.solver.get_total_partition(p) for p in range(self.molecules.get_n_polymer_types())]
            energy_total = self.hamiltonian(w_aux, total_partitions)

            phi_total = np.zeros(M)
            for i in range(S):
                phi_total += phi[self.monomer_types[i]]
            phi_diff = phi_total - 1.0

            # w_diff_i = sum_j chi_ij phi_j - P_ij w_j, zero mean (scft.py:668-679)
            w_diff = np.zeros([S, M])
            for i in range(S):
                for j in range(S):
                    w_diff[i] += self.matrix_chi[i, j] * phi[self.monomer_types[j]] - self.matrix_p[i, j] * w[j]
            for i in range(S):
                w_diff[i] -= self.cb.integral(w_diff[i]) / self.cb.get_volume()

            old_error_level = error_level
            error_level = 0.0
            error_normal = 1.0
            for i in range(S):
                error_level += self.cb.inner_product(w_diff[i], w_diff[i])
                error_normal += self.cb.inner_product(w[i], w[i])
            error_level = np.sqrt(error_level / error_normal)
            mass_error = self.cb.integral(phi_diff) / self.cb.get_volume()

            if self.box_is_altering:
                self.solver.compute_stress()
                stress_array = np.array(self.solver.get_stress())
                error_level += np.sqrt(np.sum(stress_array ** 2))
                print("%8d %12.3E " % (scft_iter, mass_error), end=" [ ")
                for p in range(self.molecules.get_n_polymer_types()):
                    print("%13.7E " % (self.solver.get_total_partition(p)), end=" ")
                print("] %15.9f %15.7E " % (energy_total, error_level), end=" ")
                print("[", ",".join(["%10.7f" % (x) for x in self.cb.get_lx()]), "]")
            else:
                print("%8d %12.3E " % (scft_iter, mass_error), end=" [ ")
                for p in range(self.molecules.get_n_polymer_types()):
                    print("%13.7E " % (self.solver.get_total_partition(p)), end=" ")
                print("] %15.9f %15.7E " % (energy_total, error_level))

            if error_level < self.tolerance:
                break

            if self.box_is_altering:
                am_current = np.concatenate((np.reshape(w, S * M), self.cb.get_lx()))
                am_diff = np.concatenate((np.reshape(w_diff, S * M), self.scale_stress * -stress_array))
                am_new = self.field_optimizer.calculate_new_fields(am_current, am_diff, old_error_level, error_level)
                w = np.reshape(am_new[0:S * M], (S, M))
                # |dL/L| at most 10 % per iteration (scft.py:729-734)
                old_lx = np.array(self.cb.get_lx())
                new_lx = np.array(am_new[-self.cb.get_dim():])
                new_dlx = np.clip((new_lx - old_lx) / old_lx, -0.1, 0.1)
                self.cb.set_lx((1 + new_dlx) * old_lx)
                self.solver.update_laplacian_operator()
            else:
                w = self.field_optimizer.calculate_new_fields(np.reshape(w, S * M), np.reshape(w_diff, S * M),
                                                              old_error_level, error_level)
                w = np.reshape(w, (S, M))
            for i in range(S):
                w[i] -= self.cb.integral(w[i]) / self.cb.get_volume()

        print("Free energy per chain (for each chain type):")
        for p in range(self.molecules.get_n_polymer_types()):
            per_chain = energy_total * self.molecules.get_polymer(p).get_alpha() / \
                self.molecules.get_polymer(p).get_volume_fraction()
            print("\tβF/n_%d : %12.7f" % (p + 1, per_chain))
        self.phi = phi
        self.w = w
