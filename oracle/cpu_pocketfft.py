"""CPU baseline in the structure of the reference's CPU platform (test and bench
infrastructure only: bench.py's cpu_baseline leg and tests/ import it; the product
never does).

The reference's CPU path runs the propagators of one schedule slice as parallel
OpenMP jobs, each job advancing its propagator with sequential MKL FFTs
(CpuComputationContinuous.cpp:148-369, :185 `#pragma omp parallel for`;
CpuComputationDiscrete.cpp:192-582; MklFFT3D.cpp:7-79 with one DFTI descriptor per
job), and then the blocks' concentrations in parallel (CpuComputationContinuous.cpp:
371-465). This module keeps that structure with numpy arithmetic and scipy.fft
(pocketfft, one worker per FFT) in a thread pool over the jobs of a slice; numpy and
pocketfft release the GIL, so the jobs run on separate cores. MKL itself has no
headers in this image, so it is not used.

Schedule: the propagators are grouped by height (a propagator's dependencies are all
of lower height), and the propagators of one height are the jobs of one slice, each
advanced over its whole contour. For the molecules of the benchmark (diblocks, the
3-arm star with a homopolymer) this is the concurrency the reference's scheduler
gives with 4 or more streams.

Arithmetic (FP64) as the reference: continuous 4th-order Richardson steps
(CpuSolverPseudo.cpp:120-188), discrete bond steps and half-bond junction steps
(:189-248), Q (CpuComputationContinuous.cpp:340-369, CpuComputationDiscrete.cpp:
563-575), Simpson / discrete contour sums for the concentrations (:371-465,
CpuComputationDiscrete.cpp:584-679) and the continuous stress (CpuSolverPseudo.cpp:
250-356, CpuComputationContinuous.cpp:643-726). Aggregated ('[') keys, q_init,
masks and solvents are not needed by the benchmarks and are refused.
"""
import concurrent.futures as cf
import os

import numpy as np
import scipy.fft

from oracle import oracle as O


def _heights(props):
    h = {}

    def height(k):
        if k not in h:
            deps = props[k]["deps"]
            h[k] = 0 if not deps else 1 + max(height(d[0]) for d in deps)
        return h[k]

    for k in props:
        height(k)
    return h


class PocketSolver:
    def __init__(self, nx, lx, plan, ds, bonds, threads=None):
        self.nx, self.lx = list(nx), list(lx)
        self.M = int(np.prod(nx))
        self.V = float(np.prod(lx))
        self.dv = self.V / self.M
        self.ds = ds
        self.model = plan["model"]
        self.plan = plan
        self.props = {p["key"]: p for p in plan["propagators"]}
        for k in self.props:
            if k.startswith("[") or k.startswith("{"):
                raise ValueError("aggregated keys and q_init are not part of this baseline")
        self.threads = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        h = _heights(self.props)
        self.levels = [sorted(k for k in self.props if h[k] == lv) for lv in range(max(h.values()) + 1)]
        self.bonds = dict(bonds)
        self._tables()

    # ---------------------------------------------------------------- tables
    def _tables(self):
        """Boltzmann bond factors (Pseudo.cpp:17-82) and weighted Fourier bases (:84-153)."""
        nx, lx = self.nx, self.lx
        k2 = []
        for d, n in enumerate(nx):
            i = np.arange(n // 2 + 1 if d == len(nx) - 1 else n)
            it = np.where(i > n // 2, n - i, i)
            k2.append((2 * np.pi * it / lx[d]) ** 2)
        grids = np.meshgrid(*k2, indexing="ij")
        ksq = sum(grids)
        self.bond = {}
        for m, b in self.bonds.items():
            self.bond[m] = np.exp(-b * b * self.ds / 6.0 * ksq)
            self.bond[m + "/2"] = np.exp(-b * b / 2 * self.ds / 6.0 * ksq)
        nz = nx[-1]
        kz = np.arange(nz // 2 + 1)
        weight = np.where((kz == 0) | (2 * kz == nz), 1.0, 2.0)
        self.basis = [g * weight for g in grids]

    def _fwd(self, q):
        return scipy.fft.rfftn(q.reshape(self.nx), workers=1)

    def _bwd(self, c):
        return scipy.fft.irfftn(c, s=self.nx, workers=1).reshape(-1)

    # ---------------------------------------------------------------- propagators
    def _advance(self, key):
        p = self.props[key]
        m = p["monomer"]
        n_max = p["max_n_segment"]
        B, Bh = self.bond[m], self.bond[m + "/2"]
        out = [None] * (n_max + 1)
        half = {}
        if self.model == "continuous":
            E, Eh = self.E[m]
            if not p["deps"]:
                out[0] = np.ones(self.M)
            else:
                q = np.ones(self.M)
                for dk, dn, rep in p["deps"]:
                    q = q * self.q[dk][dn] ** rep
                out[0] = q
            for n in range(n_max):
                q = out[n]
                q1 = E * self._bwd(B * self._fwd(E * q))
                t = E * self._bwd(Bh * self._fwd(Eh * q))
                q2 = Eh * self._bwd(Bh * self._fwd(t))
                out[n + 1] = (4.0 * q2 - q1) / 3.0
        else:
            E = self.E[m]
            if not p["deps"]:
                out[1] = E.copy()
            else:
                q = np.ones(self.M)
                for dk, dn, rep in p["deps"]:
                    q = q * self.half[dk][dn] ** rep
                out[1] = E * self._bwd(Bh * self._fwd(q))
            for n in range(1, n_max):
                out[n + 1] = E * self._bwd(B * self._fwd(out[n]))
            for n in p["junction_ends"]:
                half[n] = self._bwd(Bh * self._fwd(out[n]))
        return key, out, half

    def compute_propagators(self, w):
        if self.model == "continuous":
            self.E = {m: (np.exp(-w[m] * self.ds / 2), np.exp(-w[m] * self.ds / 4)) for m in self.bonds}
        else:
            self.E = {m: np.exp(-w[m] * self.ds) for m in self.bonds}
        self.q, self.half = {}, {}
        with cf.ThreadPoolExecutor(max_workers=self.threads) as ex:
            for level in self.levels:
                for key, out, half in ex.map(self._advance, level):
                    self.q[key] = out
                    self.half[key] = half
        self._partitions()

    def _partitions(self):
        self.Q = [None] * len(self.plan["polymers"])
        for b in self.plan["blocks"]:
            p = b["polymer"]
            if self.Q[p] is not None:
                continue
            n_agg = len(b["v_u"]) / b["n_repeated"]
            ql = self.q[b["left"]][b["n_left"]]
            if self.model == "continuous":
                qr = self.q[b["right"]][0]
            else:
                qr = self.q[b["right"]][1] / self.E[b["monomer"]]
            self.Q[p] = float(np.sum(ql * qr) * self.dv / n_agg / self.V)

    # ---------------------------------------------------------------- concentrations
    def _block_phi(self, b):
        pol = self.plan["polymers"][b["polymer"]]
        q1, q2 = self.q[b["left"]], self.q[b["right"]]
        nl, nr = b["n_left"], b["n_right"]
        norm = self.ds * pol["volume_fraction"] / pol["alpha"] / self.Q[b["polymer"]] * b["n_repeated"]
        acc = np.zeros(self.M)
        if nr == 0:
            return acc
        if self.model == "continuous":
            s = O.simpson(nr)
            for n in range(nr + 1):
                acc += s[n] * q1[nl - n] * q2[n]
            return norm * acc
        for n in range(1, nr + 1):
            acc += q1[nl - n + 1] * q2[n]
        return norm * acc / self.E[b["monomer"]]

    def compute_concentrations(self):
        with cf.ThreadPoolExecutor(max_workers=self.threads) as ex:
            phis = list(ex.map(self._block_phi, self.plan["blocks"]))
        self.phi = {}
        for b, f in zip(self.plan["blocks"], phis):
            self.phi[b["monomer"]] = self.phi.get(b["monomer"], 0.0) + f

    def get_total_concentration(self, m):
        return self.phi[m]

    # ---------------------------------------------------------------- stress (continuous)
    def _block_stress(self, b):
        q1, q2 = self.q[b["left"]], self.q[b["right"]]
        nl, nr = b["n_left"], b["n_right"]
        bsq = self.bonds[b["monomer"]] ** 2
        s = O.simpson(nr)
        out = np.zeros(len(self.nx))
        for n in range(nr + 1):
            c = bsq * np.real(self._fwd(q1[nl - n]) * np.conj(self._fwd(q2[n])))
            for d in range(len(self.nx)):
                out[d] += s[n] * float(np.sum(c * self.basis[d])) * b["n_repeated"]
        return b["polymer"], out

    def compute_stress(self):
        if self.model != "continuous":
            raise ValueError("the baseline times the continuous stress only")
        dq = np.zeros((len(self.plan["polymers"]), len(self.nx)))
        with cf.ThreadPoolExecutor(max_workers=self.threads) as ex:
            for p, v in ex.map(self._block_stress, self.plan["blocks"]):
                dq[p] += v
        stress = np.zeros(len(self.nx))
        for p, pol in enumerate(self.plan["polymers"]):
            for d in range(len(self.nx)):
                dq_dl = dq[p, d] / (-3.0 * self.lx[d] * self.M ** 2 / self.ds)
                stress[d] += dq_dl * pol["volume_fraction"] / pol["alpha"] / self.Q[p]
        return stress
