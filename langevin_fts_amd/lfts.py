"""Langevin field-theoretic simulation (L-FTS) on the hip platform.

The driver of examples/fts/lfts.py with the same parameters dictionary, the same
methods (`run`, `continue_run`, `find_saddle_point`, `compute_concentrations`,
`save_simulation_data`), the same printed rows and the same `.mat` files. The
saddle-point iteration and the Leimkuhler-Matthews update run on the GPU through
`DeviceSaddlePoint`: the auxiliary fields, concentrations and noise arrays stay in
HBM for the whole run, and only the per-iteration error levels, the Anderson-mixing
dot products and the recorded data come back to the host.

The Langevin noise comes from numpy's PCG64 generator exactly as in the reference
(lfts.py:503-507, :702), so a run with the same seed draws the same noise.
"""
import copy
import itertools
import os
import pathlib
import re
import time

import numpy as np

from . import langevinfts as lf


def calculate_sigma(langevin_nbar, langevin_dt, n_grids, volume):
    """Noise amplitude of the Langevin step (lfts.py:17-18)."""
    return np.sqrt(2 * langevin_dt * n_grids / (volume * np.sqrt(langevin_nbar)))


class SymmetricPolymerTheory:
    """Auxiliary-field representation of a multi-monomer melt (lfts.py:20-262).

    The eigenvectors of the projected chi matrix P chi P map monomer fields to
    auxiliary fields; fields along negative eigenvalues are real (fluctuating), along
    positive ones and the pressure field imaginary (found by the saddle point).
    """

    def __init__(self, monomer_types, chi_n):
        self.monomer_types = list(monomer_types)
        S = len(self.monomer_types)
        self.matrix_q = np.ones((S, S)) / S
        self.matrix_p = np.identity(S) - self.matrix_q

        eigenvalues, matrix_o = self.compute_eigen_system(chi_n, self.matrix_p)
        self.matrix_chi = self._chi_matrix(chi_n)
        self.vector_s = self.matrix_chi @ np.ones(S) / S
        self.vector_large_s = matrix_o.T @ self.vector_s

        self.aux_fields_real_idx = []
        self.aux_fields_imag_idx = []
        for i in range(S - 1):
            if np.isclose(eigenvalues[i], 0.0):
                print("One of eigenvalues is zero for given chiN values.")
            elif eigenvalues[i] > 0:
                self.aux_fields_imag_idx.append(i)
            else:
                self.aux_fields_real_idx.append(i)
        self.aux_fields_imag_idx.append(S - 1)   # the pressure field
        self.R = len(self.aux_fields_real_idx)
        self.I = len(self.aux_fields_imag_idx)
        if self.I > 1:
            print("(Warning!) For a given χN interaction parameter set, at least one of the auxiliary fields is "
                  "an imaginary field. The field fluctuations would not be fully reflected. Run this simulation "
                  "at your own risk.")

        h_const, h_coef_mu1, h_coef_mu2 = self.compute_h_coef(chi_n, eigenvalues)
        matrix_a = matrix_o.copy()
        matrix_a_inv = matrix_o.T.copy() / S
        assert np.isclose(np.std(matrix_a @ matrix_a_inv - np.identity(S)), 0.0), \
            "Invalid inverse of matrix A. Perhaps matrix O is not orthogonal."

        # d(coefficients)/d(chi N) by central differences of the eigenvalues
        # (lfts.py:84-105; vector_s is held fixed, as there)
        eps = 1e-5
        self.h_const_deriv_chin, self.h_coef_mu1_deriv_chin, self.h_coef_mu2_deriv_chin = {}, {}, {}
        for key in chi_n:
            plus, minus = dict(chi_n), dict(chi_n)
            plus[key] += eps
            minus[key] -= eps
            cp = self.compute_h_coef(plus, self.compute_eigen_system(plus, self.matrix_p)[0])
            cm = self.compute_h_coef(minus, self.compute_eigen_system(minus, self.matrix_p)[0])
            self.h_const_deriv_chin[key] = (cp[0] - cm[0]) / (2 * eps)
            self.h_coef_mu1_deriv_chin[key] = (cp[1] - cm[1]) / (2 * eps)
            self.h_coef_mu2_deriv_chin[key] = (cp[2] - cm[2]) / (2 * eps)

        self.h_const, self.h_coef_mu1, self.h_coef_mu2 = h_const, h_coef_mu1, h_coef_mu2
        self.eigenvalues = eigenvalues
        self.matrix_o = matrix_o
        self.matrix_a = matrix_a
        self.matrix_a_inv = matrix_a_inv

        print("------------ Polymer Field Theory for Multimonomer ------------")
        print("Eigenvalues:\n\t", self.eigenvalues)
        print("Eigenvectors [v1, v2, ...] :\n\t", str(self.matrix_o).replace("\n", "\n\t"))
        print("Mapping matrix A:\n\t", str(self.matrix_a).replace("\n", "\n\t"))
        print("Real Fields: ", self.aux_fields_real_idx)
        print("Imaginary Fields: ", self.aux_fields_imag_idx)
        print("In Hamiltonian:")
        print("\treference energy: ", self.h_const)
        print("\tcoefficients of int of mu(r)/V: ", self.h_coef_mu1)
        print("\tcoefficients of int of mu(r)^2/V: ", self.h_coef_mu2)
        print("\tdH_ref/dχN: ", self.h_const_deriv_chin)
        print("\td(coef of mu(r))/dχN: ", self.h_coef_mu1_deriv_chin)
        print("\td(coef of mu(r)^2)/dχN: ", self.h_coef_mu2_deriv_chin)

    def _chi_matrix(self, chi_n):
        S = len(self.monomer_types)
        chi = np.zeros((S, S))
        for i, j in itertools.combinations(range(S), 2):
            key = ",".join(sorted([self.monomer_types[i], self.monomer_types[j]]))
            if key in chi_n:
                chi[i, j] = chi[j, i] = chi_n[key]
        return chi

    def to_aux_fields(self, w):
        return self.matrix_a_inv @ w

    def to_monomer_fields(self, w_aux):
        return self.matrix_a @ w_aux

    def compute_eigen_system(self, chi_n, matrix_p):
        """Eigenpairs of P chi P ordered by |eigenvalue| descending, the last vector
        replaced by (1, ..., 1)/sqrt(S), null-space vectors orthonormalised from the
        back, first components made positive, all scaled by sqrt(S) (lfts.py:138-180)."""
        S = matrix_p.shape[0]
        vals, vecs = np.linalg.eigh(matrix_p @ self._chi_matrix(chi_n) @ matrix_p)
        order = np.argsort(np.abs(vals))[::-1]
        vals, vecs = vals[order], vecs[:, order]
        vecs[:, -1] = np.ones(S) / np.sqrt(S)
        null = np.isclose(vals, 0.0, atol=1e-12)
        vals[null] = 0.0
        basis = vecs[:, null]
        for i in range(basis.shape[1] - 2, -1, -1):
            original = basis[:, i].copy()
            for j in range(i + 1, basis.shape[1]):
                basis[:, i] -= basis[:, j] * np.dot(original, basis[:, j])
            basis[:, i] /= np.linalg.norm(basis[:, i])
        vecs[:, null] = basis
        vecs[:, vecs[0] < 0.0] *= -1.0
        return vals, vecs * np.sqrt(S)

    def compute_h_coef(self, chi_n, eigenvalues):
        """Reference energy and the coefficients of <mu> and <mu^2> (lfts.py:182-213)."""
        S = len(self.monomer_types)
        h_const = 0.5 * np.sum(self.vector_s) / S
        h_coef_mu1 = np.zeros(S - 1)
        h_coef_mu2 = np.zeros(S - 1)
        for i in range(S - 1):
            if not np.isclose(eigenvalues[i], 0.0):
                h_const -= 0.5 * self.vector_large_s[i] ** 2 / eigenvalues[i] / S
                h_coef_mu1[i] = self.vector_large_s[i] / eigenvalues[i]
                h_coef_mu2[i] = -0.5 / eigenvalues[i] * S
        return h_const, h_coef_mu1, h_coef_mu2

    def compute_hamiltonian(self, molecules, w_aux, total_partitions):
        S = len(self.monomer_types)
        h = -np.mean(w_aux[S - 1])
        for i in range(S - 1):
            h += self.h_coef_mu2[i] * np.mean(w_aux[i] ** 2)
            h += self.h_coef_mu1[i] * np.mean(w_aux[i])
        hp = 0.0
        for p in range(molecules.get_n_polymer_types()):
            pol = molecules.get_polymer(p)
            hp -= pol.get_volume_fraction() / pol.get_alpha() * np.log(total_partitions[p])
        return hp + h + self.h_const

    def func_deriv_coefficients(self, indices):
        """dH/dw_aux[i] = cw . w_aux + cp . phi + c0 for each i of `indices`, the sign
        of the imaginary fields applied (lfts.py:219-245)."""
        S = len(self.monomer_types)
        n = len(indices)
        cw, cp, c0 = np.zeros((n, S)), np.zeros((n, S)), np.zeros(n)
        for r, i in enumerate(indices):
            if i != S - 1:
                cw[r, i] = 2 * self.h_coef_mu2[i]
                c0[r] = self.h_coef_mu1[i]
                cp[r] = self.matrix_a[:, i]
            else:
                cp[r] = 1.0
                c0[r] = -1.0
            if i in self.aux_fields_imag_idx:
                cw[r], cp[r], c0[r] = -cw[r], -cp[r], -c0[r]
        return cw, cp, c0

    def compute_func_deriv(self, w_aux, phi, indices):
        cw, cp, c0 = self.func_deriv_coefficients(indices)
        phi_rows = np.array([phi[t] for t in self.monomer_types])
        h_deriv = cw @ w_aux + cp @ phi_rows + c0[:, None]
        return h_deriv, {}

    def compute_h_deriv_chin(self, chi_n, w_aux=None, moments=None):
        """dH/d(chi N) (lfts.py:248-262); `moments` = (<w_aux>, <w_aux^2>) replaces w_aux."""
        S = len(self.monomer_types)
        if moments is None:
            moments = ([np.mean(w_aux[i]) for i in range(S)], [np.mean(w_aux[i] ** 2) for i in range(S)])
        mean, mean_sq = moments
        dH = {}
        for key in chi_n:
            dH[key] = self.h_const_deriv_chin[key]
            for i in range(S - 1):
                dH[key] += self.h_coef_mu2_deriv_chin[key][i] * mean_sq[i]
                dH[key] += self.h_coef_mu1_deriv_chin[key][i] * mean[i]
        return dH


def guess_initial_params(seed, n_in=1):
    """He-normal weights and zero biases for the learned-guess network (layout of
    LearnedGuess: W0[n_in][27][32], b0[32], 4 x (W[27][32][32], b[32]), WL[27][32], bL)."""
    C, T = 32, 27
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0.0, np.sqrt(2.0 / (T * n_in)), n_in * T * C), np.zeros(C)]
    for _ in range(4):
        parts += [rng.normal(0.0, np.sqrt(2.0 / (T * C)), T * C * C), np.zeros(C)]
    parts += [rng.normal(0.0, np.sqrt(1.0 / (T * C)), T * C), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


GUESS_MODES = {"absolute": 0, "delta": 1, "residual": 2}
# training steps whose samples' mean deviations become the network's scales
GUESS_SCALE_STEPS = 8
# training steps whose mean loss decides whether the residual guess is used while it trains
GUESS_USE_WINDOW = 20


class NoiseStream:
    """The Langevin noise of one numpy Generator, drawn one step ahead on a helper
    thread: the next step's noise is filled while the device runs the current saddle
    point (numpy releases the GIL while it fills an array, and the saddle-point call
    releases it while it waits on the device). Draws come in the order of direct
    calls; rewind() puts the generator back to the state before a pending draw, so a
    checkpoint stores the state the next step will draw from."""

    def __init__(self, generator, sigma, shape):
        from concurrent.futures import ThreadPoolExecutor

        self.generator = generator
        self.sigma = sigma
        self.shape = list(shape)
        self.pool = ThreadPoolExecutor(max_workers=1)
        self.pending = None   # (generator state before the draw, future)

    def _launch(self):
        state = self.generator.bit_generator.state
        self.pending = (state, self.pool.submit(self.generator.normal, 0.0, self.sigma, self.shape))

    def next(self):
        if self.pending is None:
            self._launch()
        _, future = self.pending
        noise = future.result()
        self._launch()
        return noise

    def rewind(self):
        if self.pending is not None:
            state, future = self.pending
            future.result()
            self.generator.bit_generator.state = state
            self.pending = None


class LFTS:
    """L-FTS driver; `params` has the keys of the reference examples (fts/Lamella.py).
    An optional "learned_guess" entry {"train_steps", "lr", "lr_final", "train_iters",
    "replay", "inputs", "use_loss", "seed"} trains a periodic 3D CNN on the saddle points of the
    first train_steps accepted Langevin steps and then starts every saddle point from
    its prediction. "inputs": "delta" (default: a network f(w-) ~ w+, used as
    w+ += f(w-_new) - f(w-_old)), "absolute" (w+ = f(w-)) or "residual" (the first
    saddle iteration evaluates the previous w+ as plain mixing would; a network g of
    w-, that iteration's residual and w+ predicts the correction of w+, which replaces
    the first mixing step; that evaluation stays in the mixing history). The residual
    guess also starts the saddle points of the training steps once its recent training
    loss is below "use_loss" (default 0.2; 0: only after training); its training samples
    are the same either way. "train_iters": Adam steps per training step (48)."""

    def __init__(self, params, random_seed=None, n_replicas=1):
        """n_replicas > 1: that many independent runs of the same system advance together,
        their propagators batched into the same launches; replica r draws its noise from
        PCG64(random_seed + r) (replica_step)."""
        self.n_replicas = int(n_replicas)
        assert self.n_replicas >= 1, "n_replicas must be positive"
        self.monomer_types = sorted(list(params["segment_lengths"].keys()))
        self.segment_lengths = copy.deepcopy(params["segment_lengths"])
        self.distinct_polymers = copy.deepcopy(params["distinct_polymers"])
        assert len(self.monomer_types) == len(set(self.monomer_types)), "There are duplicated monomer_types"

        platform = params.get("platform", "hip")
        factory = lf.PlatformSelector.create_factory(platform, False)
        factory.display_info()
        cb = factory.create_computation_box(params["nx"], params["lx"])

        # chi N of every pair, missing pairs zero (lfts.py:305-327)
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

        # random copolymers: one block whose segment length is the fraction-weighted
        # mean square of its constituents (lfts.py:365-397)
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
            molecules.add_polymer(polymer["volume_fraction"], polymer["blocks_input"])
        analyzer = factory.create_propagator_analyzer(molecules, params.get("aggregate_propagator_computation", True))
        n_grid = int(np.prod(params["nx"]))
        am_params = [params["am"]["max_hist"], params["am"]["start_error"], params["am"]["mix_min"],
                     params["am"]["mix_init"]]
        if self.n_replicas == 1:
            solver = factory.create_pseudospectral_solver(cb, molecules, analyzer)
            am = factory.create_anderson_mixing(len(self.mpt.aux_fields_imag_idx) * n_grid, *am_params)
        else:
            solver = factory.create_replica_solver(cb, molecules, analyzer, self.n_replicas)
            am = None
        langevin_sigma = calculate_sigma(params["langevin"]["nbar"], params["langevin"]["dt"], n_grid,
                                         np.prod(params["lx"]))

        S = len(self.monomer_types)
        self.dt_scaling = np.ones(S)
        for i in range(S - 1):
            self.dt_scaling[i] = np.abs(self.mpt.eigenvalues[i]) / np.max(np.abs(self.mpt.eigenvalues))

        self.random_bg = np.random.PCG64() if random_seed is None else np.random.PCG64(random_seed)
        self.random = np.random.Generator(self.random_bg)
        self.noise_stream = None   # NoiseStream over self.random, made on the first next_noise()
        if self.n_replicas > 1:
            base = 0 if random_seed is None else random_seed
            self.replica_random = [np.random.Generator(np.random.PCG64(base + r)) for r in range(self.n_replicas)]

        # the saddle point on the device
        mpt = self.mpt
        rcw, rcp, rc0 = mpt.func_deriv_coefficients(mpt.aux_fields_real_idx)
        icw, icp, ic0 = mpt.func_deriv_coefficients(mpt.aux_fields_imag_idx)
        h_lin = np.zeros(S)
        h_sq = np.zeros(S)
        h_lin[S - 1] = -1.0
        h_lin[:S - 1] = mpt.h_coef_mu1
        h_sq[:S - 1] = mpt.h_coef_mu2
        self.saddle_device = lf.DeviceSaddlePoint(
            solver, am, self.monomer_types, {k: dict(v) for k, v in self.random_fraction.items()},
            mpt.matrix_a.ravel().tolist(), mpt.aux_fields_real_idx, mpt.aux_fields_imag_idx,
            rcw.ravel().tolist(), rcp.ravel().tolist(), rc0.tolist(), icw.ravel().tolist(), icp.ravel().tolist(),
            ic0.tolist(), self.dt_scaling.tolist(), h_lin.tolist(), h_sq.tolist(), float(mpt.h_const),
            am_params if self.n_replicas > 1 else [])
        # mixed precision (an addition): saddle iterations on an FP32 solver of the same molecules until
        # the error falls below the switch value, then FP64 iterations, which alone can meet the tolerance
        self.low_precision_solver = None
        if params["saddle"].get("mixed_precision", False):
            if platform != "hip":
                raise ValueError("saddle.mixed_precision needs the FP64 platform \"hip\"")
            f32 = lf.PlatformSelector.create_factory("hip-fp32", False)
            if self.n_replicas == 1:
                self.low_precision_solver = f32.create_pseudospectral_solver(cb, molecules, analyzer)
            else:
                self.low_precision_solver = f32.create_replica_solver(cb, molecules, analyzer, self.n_replicas)
            switch = params["saddle"].get("mixed_precision_switch", params["saddle"]["tolerance"])
            self.saddle_device.set_low_precision_solver(self.low_precision_solver, switch)

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
        for p in range(molecules.get_n_polymer_types()):
            print("distinct_polymers[%d]:" % (p))
            print("\tvolume fraction: %f, alpha: %f, N: %d" % (
                molecules.get_polymer(p).get_volume_fraction(), molecules.get_polymer(p).get_alpha(),
                molecules.get_polymer(p).get_n_segment_total()))
        print("Invariant Polymerization Index (N_Ref): %d" % (params["langevin"]["nbar"]))
        print("Langevin Sigma: %f" % (langevin_sigma))
        print("Scaling factor of delta tau N for each field: ", self.dt_scaling)
        print("Random Number Generator: ", self.random_bg.state)
        analyzer.display_blocks()
        analyzer.display_propagators()

        self.params = params
        self.chain_model = params["chain_model"]
        self.ds = params["ds"]
        self.langevin = params["langevin"].copy()
        self.langevin.update({"sigma": langevin_sigma})
        self.verbose_level = params["verbose_level"]
        # learned initial guess of w+ (optional):
        # {"train_steps", "lr", "lr_final", "train_iters", "replay", "inputs", "use_loss", "seed"}
        self.learned = dict(params["learned_guess"]) if params.get("learned_guess") else None
        if self.learned is not None:
            assert len(self.mpt.aux_fields_real_idx) == 1 and len(self.mpt.aux_fields_imag_idx) == 1, \
                "the learned guess maps one real field to one imaginary field"
            assert cb.get_dim() == 3, "the learned guess needs a 3D grid"
            self.learned.setdefault("train_steps", 1000)
            self.learned.setdefault("lr", 5e-4)
            self.learned.setdefault("lr_final", self.learned["lr"] / 10)
            self.learned.setdefault("train_iters", 48)
            self.learned.setdefault("replay", 64)
            if "inputs" not in self.learned:
                # the boolean form: "delta": True / False
                self.learned["inputs"] = "delta" if self.learned.get("delta", True) else "absolute"
            assert self.learned["inputs"] in GUESS_MODES, "inputs must be delta, absolute or residual"
            self.learned.setdefault("seed", 0)
        self.guess = None
        self.guess_trained_steps = 0
        self.guess_losses = []          # the newest sample's loss at each training step
        self.guess_scale_samples = []   # sample deviations collected before the network exists
        self.saddle = params["saddle"].copy()
        self.recording = params["recording"].copy()
        self.cb = cb
        self.molecules = molecules
        self.propagator_analyzer = analyzer
        self.solver = solver
        self.am = am

    # ---------------------------------------------------------------- host-array API
    def compute_concentrations(self, w_aux):
        """phi of every monomer type for host fields w_aux (lfts.py:534-567)."""
        S = len(self.monomer_types)
        w = self.mpt.to_monomer_fields(w_aux)
        w_input = {self.monomer_types[i]: w[i] for i in range(S)}
        for name, fractions in self.random_fraction.items():
            w_input[name] = np.zeros(self.cb.get_n_grid())
            for t, f in fractions.items():
                w_input[name] += w_input[t] * f
        t0 = time.time()
        self.solver.compute_propagators(w_input)
        elapsed = {"solver": time.time() - t0}
        t0 = time.time()
        self.solver.compute_concentrations()
        phi = {t: self.solver.get_total_concentration(t) for t in self.monomer_types}
        for name, fractions in self.random_fraction.items():
            phi[name] = self.solver.get_total_concentration(name)
            for t, f in fractions.items():
                phi[t] += phi[name] * f
        elapsed["phi"] = time.time() - t0
        return phi, elapsed

    def find_saddle_point(self, w_aux):
        """Saddle point of the imaginary fields for host w_aux, updated in place
        (lfts.py:806-867); returns (phi, hamiltonian, iterations, error level)."""
        self.saddle_device.set_fields(np.ascontiguousarray(w_aux, dtype=np.float64))
        n_iter, error_level, hamiltonian = self._find_saddle_point_device()
        w_aux[...] = self.saddle_device.get_fields()
        return self._phi_from_device(), hamiltonian, n_iter, error_level

    # ---------------------------------------------------------------- learned guess
    def guess_sample_stds(self):
        """Standard deviations of the current training sample: the network's inputs, then
        its target. One input (delta, absolute): w- and w+. Residual: w-, the first
        residual, the starting w+ and the converged w+ minus the starting w+."""
        if self.learned["inputs"] == "residual":
            return [float(np.std(row)) for row in self.saddle_device.guess_sample()]
        w_aux = self.saddle_device.get_fields()
        ri, ii = self.mpt.aux_fields_real_idx[0], self.mpt.aux_fields_imag_idx[0]
        return [float(np.std(w_aux[ri])), float(np.std(w_aux[ii]))]

    def init_guess(self, stds=None):
        """The network, with He-normal weights from the configured seed, its inputs divided
        and its output multiplied by the given standard deviations (inputs, then target;
        default: those of the current sample)."""
        stds = list(self.guess_sample_stds() if stds is None else stds)
        in_scales = [float(v) or 1.0 for v in stds[:-1]]
        out_scale = float(stds[-1]) or 1.0
        self.guess = lf.LearnedGuess(self.cb.get_nx(), in_scales, out_scale)
        self.guess.set_params(guess_initial_params(self.learned["seed"], len(in_scales)))
        return in_scales, out_scale

    def guess_usable(self):
        """Whether the residual guess may start the saddle points while it still trains: its
        newest samples' mean loss over the last GUESS_USE_WINDOW training steps is below
        "use_loss" (in the scaled units, where 1 is about the target's variance). Its
        training samples do not depend on this: the first evaluation is at the step's
        starting w+ either way, and the target is the converged w+ minus that start."""
        use_loss = self.learned.get("use_loss", 0.2)
        return (self.guess is not None and self.learned["inputs"] == "residual" and use_loss > 0 and
                len(self.guess_losses) >= GUESS_USE_WINDOW and
                float(np.mean(self.guess_losses[-GUESS_USE_WINDOW:])) < use_loss)

    def _guess_before_update(self, ready):
        if self.learned is None:
            return
        if self.learned["inputs"] == "delta" and ready:
            self.saddle_device.prepare_guess(self.guess)
        if self.learned["inputs"] == "residual":
            # record the first residual while training; apply the network once it is ready, or
            # while it trains once its loss is low enough
            use = ready or self.guess_usable()
            self.saddle_device.set_residual_guess(self.guess if use else None, not ready)

    def _guess_after_update(self, ready):
        if ready and self.learned["inputs"] != "residual":
            self.saddle_device.predict_guess(self.guess, GUESS_MODES[self.learned["inputs"]])

    def train_guess(self):
        """Adam steps on the current sample and on samples replayed from a device ring of
        recent saddle points. The network is created at the first call, scaled by that
        sample's standard deviations. The deviations of the first GUESS_SCALE_STEPS
        samples are kept. If the first sample's fall below half their mean in any input or
        the target, the network is created again with those raised to that floor, and
        trains on from the replay ring. After a random start the fields grow for many
        steps (w+ 0.014 in the first sample, 3 twenty steps later at 64^3), and scales
        that small have left the network dead. The learning rate falls from "lr" to
        "lr_final" on a cosine over the "train_steps" training steps. Returns the newest
        sample's loss."""
        if len(self.guess_scale_samples) < GUESS_SCALE_STEPS:
            self.guess_scale_samples.append(self.guess_sample_stds())
            if self.guess is None:
                self.init_guess(self.guess_scale_samples[0])
            elif len(self.guess_scale_samples) == GUESS_SCALE_STEPS:
                first = np.asarray(self.guess_scale_samples[0])
                floor = 0.5 * np.mean(self.guess_scale_samples, axis=0)
                if np.any(first < floor):
                    self.init_guess(np.maximum(first, floor))
        lr0, lr1 = self.learned["lr"], self.learned["lr_final"]
        frac = min(self.guess_trained_steps / max(self.learned["train_steps"], 1), 1.0)
        lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + np.cos(np.pi * frac))
        loss = self.saddle_device.train_guess(self.guess, lr, self.learned["train_iters"],
                                              self.learned["replay"])
        self.guess_losses.append(loss)
        self.guess_trained_steps += 1
        return loss

    def next_noise(self):
        """The next step's Langevin noise N(0, sigma^2) of shape [R, M] from self.random,
        drawn ahead on a helper thread (NoiseStream)."""
        if self.noise_stream is None:
            R = len(self.mpt.aux_fields_real_idx)
            self.noise_stream = NoiseStream(self.random, self.langevin["sigma"], [R, self.cb.get_n_grid()])
        return self.noise_stream.next()

    def rewind_noise(self):
        """Undo a draw made ahead, so self.random's state is that of the next step."""
        if self.noise_stream is not None:
            self.noise_stream.rewind()

    def guess_ready(self):
        return self.guess is not None and self.guess_trained_steps >= self.learned["train_steps"]

    def replica_step(self):
        """One Langevin step of every replica (n_replicas > 1): per-replica noise, one
        batched saddle-point iteration over all replicas; replicas whose saddle point
        missed the tolerance are put back to their previous state (their step is
        skipped). Returns (iterations, error levels, accepted flags) per replica."""
        assert self.n_replicas > 1, "replica_step needs n_replicas > 1"
        R = len(self.mpt.aux_fields_real_idx)
        M = self.cb.get_n_grid()
        dev = self.saddle_device
        dev.backup()
        noise = np.stack([g.normal(0.0, self.langevin["sigma"], [R, M]) for g in self.replica_random])
        dev.langevin_update(noise, self.langevin["dt"])
        its, errs, _, _ = dev.find_saddle_point(self.saddle["tolerance"], self.saddle["max_iter"], 0)
        ok = [not (np.isnan(e) or e >= self.saddle["tolerance"]) for e in errs]
        if not all(ok):
            dev.restore_replicas([0 if a else 1 for a in ok])
        return its, errs, ok

    def learned_step(self, noise):
        """One Langevin step with the learned guess when it is trained, else with training
        (the body of run's loop without recording); returns (iterations, error, accepted)."""
        dev = self.saddle_device
        dev.backup()
        ready = self.guess_ready()
        self._guess_before_update(ready)
        dev.langevin_update(noise, self.langevin["dt"])
        self._guess_after_update(ready)
        n, err, _ = self._find_saddle_point_device()
        ok = not (np.isnan(err) or err >= self.saddle["tolerance"])
        if not ok:
            dev.restore()
        elif not ready:
            self.train_guess()
        return n, err, ok

    # ---------------------------------------------------------------- device-resident run
    def _phi_from_device(self):
        return {t: self.saddle_device.get_concentration(i) for i, t in enumerate(self.monomer_types)}

    def _find_saddle_point_device(self):
        verbose = self.verbose_level if self.verbose_level in (1, 2) else 0
        n_iter, error_level, hamiltonian, log = self.saddle_device.find_saddle_point(
            self.saddle["tolerance"], self.saddle["max_iter"], verbose)
        P = self.molecules.get_n_polymer_types()
        rows = log[:n_iter] if verbose == 2 else (log if verbose == 1 else [])
        for r in rows:
            print("%8d %12.3E " % (int(r[0]), r[1]), end=" [ ")
            for p in range(P):
                print("%13.7E " % (r[3 + p]), end=" ")
            print("] %15.9f   [" % (r[2]), end="")
            for e in r[3 + P:]:
                print("%13.7E" % (e), end=" ")
            print("]")
        return n_iter, error_level, hamiltonian

    def structure_function_terms(self, w_aux, phi):
        """One sample of the structure functions, S_ij(k) += mu_i(k) conj(phi_j(k)) for
        every monomer pair i <= j (lfts.py:753-766): mu_i(k) = sum over the real-field
        eigenmodes k < S-1 of FFT(w_aux[k]) A_inv[k, i] / eigenvalue_k / M and
        phi_j(k) = FFT(phi_j) / M, on the half spectrum of the grid."""
        S = len(self.monomer_types)
        nx = self.cb.get_nx()
        M = self.cb.get_n_grid()
        mu_k, phi_k = {}, {}
        w_aux_k = [np.fft.rfftn(np.reshape(w_aux[k], nx)) for k in range(S - 1)]
        for i, t in enumerate(self.monomer_types):
            phi_k[t] = np.fft.rfftn(np.reshape(phi[t], nx)) / M
            mu_k[t] = np.zeros_like(phi_k[t], np.complex128)
            for k in range(S - 1):
                mu_k[t] += w_aux_k[k] * self.mpt.matrix_a_inv[k, i] / self.mpt.eigenvalues[k] / M
        out = {}
        for a, b in itertools.combinations_with_replacement(range(S), 2):
            key = self.monomer_types[a] + "," + self.monomer_types[b]
            x, y = sorted(key.split(","))
            out[key] = mu_k[x] * np.conj(phi_k[y])
        return out

    def save_simulation_data(self, path, w, phi, langevin_step, normal_noise_prev):
        """`.mat` checkpoint with the keys of lfts.py:569-598."""
        self.rewind_noise()   # the stored state is the one the next step draws from
        from scipy.io import savemat

        mdic = {
            "initial_params": self.params,
            "dim": self.cb.get_dim(), "nx": self.cb.get_nx(), "lx": self.cb.get_lx(),
            "monomer_types": self.monomer_types, "chi_n": dict(self.chi_n), "chain_model": self.chain_model,
            "ds": self.ds, "dt": self.langevin["dt"], "nbar": self.langevin["nbar"],
            "eigenvalues": self.mpt.eigenvalues,
            "aux_fields_real": self.mpt.aux_fields_real_idx,
            "aux_fields_imag": self.mpt.aux_fields_imag_idx,
            "matrix_a": self.mpt.matrix_a, "matrix_a_inverse": self.mpt.matrix_a_inv,
            "langevin_step": langevin_step,
            "random_generator": self.random_bg.state["bit_generator"],
            "random_state_state": str(self.random_bg.state["state"]["state"]),
            "random_state_inc": str(self.random_bg.state["state"]["inc"]),
            "normal_noise_prev": normal_noise_prev}
        for i, name in enumerate(self.monomer_types):
            mdic["w_" + name] = w[i]
        for name in self.monomer_types:
            mdic["phi_" + name] = phi[name]
        savemat(path, mdic, long_field_names=True, do_compression=True)

    def continue_run(self, file_name):
        """Resume from a checkpoint with the stored PCG64 state (lfts.py:600-629)."""
        from scipy.io import loadmat

        data = loadmat(file_name, squeeze_me=True)
        period = self.recording["sf_recording_period"]
        if data["langevin_step"] % period != 0:
            print(f"(Warning!) 'langevin_step' of {file_name} is not a multiple of 'sf_recording_period'.")
            nxt = (data["langevin_step"] // period + 1) * period
            print(f"The structure function will be correctly recorded after {nxt}th langevin_step.")
        self.rewind_noise()
        self.random_bg.state = {"bit_generator": "PCG64",
                                "state": {"state": int(data["random_state_state"]),
                                          "inc": int(data["random_state_inc"])},
                                "has_uint32": 0, "uinteger": 0}
        print("Restored Random Number Generator: ", self.random_bg.state)
        initial_fields = {name: np.array(data["w_" + name]) for name in self.monomer_types}
        self.run(initial_fields=initial_fields, normal_noise_prev=data["normal_noise_prev"],
                 start_langevin_step=data["langevin_step"] + 1)

    def run(self, initial_fields, normal_noise_prev=None, start_langevin_step=None):
        """Langevin dynamics of the real fields with a saddle point of the imaginary
        fields at every step (lfts.py:631-804)."""
        print("---------- Run  ----------")
        S = len(self.monomer_types)
        R = len(self.mpt.aux_fields_real_idx)
        M = self.cb.get_n_grid()
        nx = self.cb.get_nx()
        dev = self.saddle_device
        pathlib.Path(self.recording["dir"]).mkdir(parents=True, exist_ok=True)

        w = np.zeros([S, M])
        for i in range(S):
            w[i] = np.reshape(initial_fields[self.monomer_types[i]], M)
        w_aux = self.mpt.to_aux_fields(w)
        dev.set_fields(np.ascontiguousarray(w_aux))

        print("iterations, mass error, total partitions, Hamiltonian, incompressibility error (or saddle point error)")
        self._find_saddle_point_device()

        H_history = []
        dH_history = {key: [] for key in self.chi_n}
        sf_average = {}
        for a, b in itertools.combinations_with_replacement(range(S), 2):
            key = self.monomer_types[a] + "," + self.monomer_types[b]
            sf_average[key] = np.zeros(tuple(nx[:-1]) + (nx[-1] // 2 + 1,), np.complex128)

        if normal_noise_prev is None:
            normal_noise_prev = np.zeros([R, M])
        dev.set_noise_prev(np.ascontiguousarray(np.reshape(normal_noise_prev, [R, M]), dtype=np.float64))
        if start_langevin_step is None:
            start_langevin_step = 1

        saddle_fail_count = 0
        successive_fail_count = 0
        total_saddle_iter = 0
        total_error_level = 0.0
        time_start = time.time()
        hamiltonian = None
        for langevin_step in range(start_langevin_step, self.langevin["max_step"] + 1):
            print("Langevin step: ", langevin_step)
            dev.backup()
            noise = self.next_noise()
            ready = self.guess_ready()
            self._guess_before_update(ready)
            dev.langevin_update(noise, self.langevin["dt"])
            self._guess_after_update(ready)
            saddle_iter, error_level, hamiltonian = self._find_saddle_point_device()
            total_saddle_iter += saddle_iter
            total_error_level += error_level

            if np.isnan(error_level) or error_level >= self.saddle["tolerance"]:
                if successive_fail_count < 5:
                    print("The tolerance of the saddle point was not met. Langevin random noise is regenerated.")
                    dev.restore()
                    successive_fail_count += 1
                    saddle_fail_count += 1
                    continue
                print("The tolerance of the saddle point was not met %d times in a row. Simulation is aborted."
                      % (successive_fail_count))
                break
            successive_fail_count = 0
            if self.learned is not None and not self.guess_ready():
                self.train_guess()

            computing = langevin_step % self.recording["sf_computing_period"] == 0
            recording_sf = langevin_step % self.recording["sf_recording_period"] == 0
            if computing:
                H_history.append(hamiltonian)
                dH = self.mpt.compute_h_deriv_chin(self.chi_n, moments=dev.field_moments())
                for key in self.chi_n:
                    dH_history[key].append(dH[key])
            if recording_sf:
                from scipy.io import savemat

                mdic = {"H_history": np.array(H_history)}
                for key in self.chi_n:
                    pair = sorted(key.split(","))
                    mdic["dH_history_" + pair[0] + "_" + pair[1]] = np.array(dH_history[key])
                savemat(os.path.join(self.recording["dir"], "dH_%06d.mat" % (langevin_step)), mdic,
                        long_field_names=True, do_compression=True)
                H_history = []
                dH_history = {key: [] for key in self.chi_n}

            if computing:
                terms = self.structure_function_terms(dev.get_fields(), self._phi_from_device())
                for key in sf_average:
                    sf_average[key] += terms[key]
            if recording_sf:
                from scipy.io import savemat

                chi_n_mat = {}
                for key in self.chi_n:
                    pair = sorted(key.split(","))
                    chi_n_mat[pair[0] + "," + pair[1]] = self.chi_n[key]
                mdic = {"dim": self.cb.get_dim(), "nx": self.cb.get_nx(), "lx": self.cb.get_lx(),
                        "chi_n": chi_n_mat, "chain_model": self.chain_model, "ds": self.ds,
                        "dt": self.langevin["dt"], "nbar": self.langevin["nbar"], "initial_params": self.params}
                for key in sf_average:
                    sf_average[key] *= self.recording["sf_computing_period"] / self.recording["sf_recording_period"] \
                        * self.cb.get_volume() * np.sqrt(self.langevin["nbar"])
                    pair = sorted(key.split(","))
                    mdic["structure_function_" + pair[0] + "_" + pair[1]] = sf_average[key]
                savemat(os.path.join(self.recording["dir"], "structure_function_%06d.mat" % (langevin_step)), mdic,
                        long_field_names=True, do_compression=True)
                for key in sf_average:
                    sf_average[key][...] = 0.0

            if langevin_step % self.recording["recording_period"] == 0:
                w_aux = dev.get_fields()
                self.save_simulation_data(
                    path=os.path.join(self.recording["dir"], "fields_%06d.mat" % (langevin_step)),
                    w=self.mpt.to_monomer_fields(w_aux), phi=self._phi_from_device(), langevin_step=langevin_step,
                    normal_noise_prev=dev.get_noise_prev())

        print("The number of times that tolerance of saddle point was not met and Langevin random noise was "
              "regenerated: %d times" % (saddle_fail_count))
        n_steps = self.langevin["max_step"] + 1 - start_langevin_step
        duration = time.time() - time_start
        return total_saddle_iter, total_saddle_iter / n_steps, duration / n_steps, total_error_level / n_steps
