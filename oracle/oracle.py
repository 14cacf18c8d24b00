"""ctypes front end of the C oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product never imports it. The plan argument has the shape that
PropagatorAnalyzer's propagators/blocks take (PropagatorAnalyzer.cpp:142-269):

    {"model": "continuous"|"discrete",
     "polymers": [{"alpha", "volume_fraction", "blocks": [{"v", "u", "key_vu", "key_uv", ...}]}],
     "propagators": [{"key", "max_n_segment", "monomer", "deps": [[key, n, r]], "junction_ends"}],
     "blocks": [{"polymer", "left", "right", "monomer", "n_left", "n_right", "n_repeated", "v_u"}]}

in ComparePropagatorKey / std::map order.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -f oracle/Makefile)")
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_grid_create.restype = vp
        L.orc_grid_create.argtypes = [i, ctypes.POINTER(i), dp, dp]
        L.orc_grid_destroy.argtypes = [vp]
        L.orc_grid_set_lx.argtypes = [vp, dp]
        L.orc_grid_volume.restype = d
        L.orc_grid_volume.argtypes = [vp]
        L.orc_grid_nc.restype = i
        L.orc_grid_nc.argtypes = [vp]
        L.orc_integral.restype = d
        L.orc_integral.argtypes = [vp, dp]
        L.orc_inner_product.restype = d
        L.orc_inner_product.argtypes = [vp, dp, dp]
        L.orc_fft_forward.argtypes = [vp, dp, dp]
        L.orc_fft_backward.argtypes = [vp, dp, dp]
        L.orc_boltz_bond.argtypes = [vp, d, d, dp]
        L.orc_fourier_basis.argtypes = [vp, dp, dp, dp]
        L.orc_solver_create.restype = vp
        L.orc_solver_create.argtypes = [vp, vp]
        L.orc_solver_destroy.argtypes = [vp]
        L.orc_compute_propagators.restype = i
        L.orc_compute_propagators.argtypes = [vp, ctypes.POINTER(dp), ctypes.POINTER(dp)]
        L.orc_compute_concentrations.argtypes = [vp]
        L.orc_compute_stress.argtypes = [vp]
        L.orc_partition.restype = d
        L.orc_partition.argtypes = [vp, i]
        L.orc_solvent_partition.restype = d
        L.orc_solvent_partition.argtypes = [vp, i]
        L.orc_block_phi.restype = dp
        L.orc_block_phi.argtypes = [vp, i]
        L.orc_solvent_phi.restype = dp
        L.orc_solvent_phi.argtypes = [vp, i]
        L.orc_dq_dl.restype = d
        L.orc_dq_dl.argtypes = [vp, i, i]
        L.orc_propagator.restype = dp
        L.orc_propagator.argtypes = [vp, i, i, i]
        L.orc_simpson.argtypes = [i, dp]
        L.orc_solver_set_realspace.restype = i
        L.orc_solver_set_realspace.argtypes = [vp, ctypes.POINTER(i)]
        L.orc_laplacian.argtypes = [vp, ctypes.POINTER(i), d, d, dp]
        L.orc_tridiagonal.argtypes = [dp, dp, dp, dp, i, dp, i]
        L.orc_tridiagonal_periodic.argtypes = [dp, dp, dp, dp, i, dp, i]
        L.orc_am_create.restype = vp
        L.orc_am_create.argtypes = [i, i, d, d, d]
        L.orc_am_destroy.argtypes = [vp]
        L.orc_am_reset.argtypes = [vp]
        L.orc_am_new_fields.argtypes = [vp, dp, dp, dp, d, d]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _Plan(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("ds", ctypes.c_double), ("n_monomers", ctypes.c_int),
                ("segment_length", ctypes.c_void_p), ("n_props", ctypes.c_int)] + \
        [(n, ctypes.c_void_p) for n in ["prop_monomer", "prop_max_n", "prop_kind", "prop_qinit",
                                        "prop_dep_off", "prop_n_deps", "dep_prop", "dep_n", "dep_rep",
                                        "prop_je_off", "prop_n_je", "je"]] + \
        [("n_blocks", ctypes.c_int)] + \
        [(n, ctypes.c_void_p) for n in ["blk_polymer", "blk_left", "blk_right", "blk_n_left",
                                        "blk_n_right", "blk_rep", "blk_monomer", "blk_n_vu"]] + \
        [("n_polymers", ctypes.c_int), ("pol_vf", ctypes.c_void_p), ("pol_alpha", ctypes.c_void_p),
         ("n_solvents", ctypes.c_int), ("sol_vf", ctypes.c_void_p), ("sol_monomer", ctypes.c_void_p)]


def q_input_name(key):
    """PropagatorCode.cpp:232-247: name inside the leading {...} of a key."""
    return key[1:key.rindex("}")]


def simpson(n):
    out = np.empty(n + 1)
    lib().orc_simpson(n, _dp(out))
    return out


BC_CODES = {"periodic": 0, "reflecting": 1, "absorbing": 2}


def _bc_array(bc, dim):
    bc = list(bc) if bc else ["periodic"] * (2 * dim)
    if len(bc) != 2 * dim:
        raise ValueError("expected 0 or %d boundary conditions" % (2 * dim))
    return (ctypes.c_int * len(bc))(*[BC_CODES[b.lower()] for b in bc])


def laplacian(grid, bc, bond_length_sq, ds):
    """FiniteDifference::get_laplacian_matrix: per axis a dict of l, d, h."""
    t = np.zeros(3 * sum(grid.nx))
    lib().orc_laplacian(grid.h, _bc_array(bc, grid.dim), bond_length_sq, ds, _dp(t))
    out, o = [], 0
    for n in grid.nx:
        out.append({"l": t[o:o + n].copy(), "d": t[o + n:o + 2 * n].copy(), "h": t[o + 2 * n:o + 3 * n].copy()})
        o += 3 * n
    return out


def tridiagonal(l, d, h, rhs, periodic=False):
    """CpuSolverReal::tridiagonal (periodic: tridiagonal_periodic)."""
    l, d, h, rhs = (np.ascontiguousarray(v, dtype=np.float64) for v in (l, d, h, rhs))
    x = np.zeros(len(rhs))
    f = lib().orc_tridiagonal_periodic if periodic else lib().orc_tridiagonal
    f(_dp(l), _dp(d), _dp(h), _dp(x), 1, _dp(rhs), len(rhs))
    return x


class Grid:
    """ComputationBox (ComputationBox.cpp:10-308) + the FP64 FFT."""

    def __init__(self, nx, lx, mask=None):
        self.nx = list(nx)
        self.dim = len(nx)
        self.M = int(np.prod(nx))
        nxa = (ctypes.c_int * self.dim)(*nx)
        lxa = np.array(lx, dtype=np.float64)
        self._mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
        self.h = lib().orc_grid_create(self.dim, nxa, _dp(lxa),
                                       None if self._mask is None else _dp(self._mask))
        self.Mc = lib().orc_grid_nc(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_destroy(self.h)
            self.h = None

    def set_lx(self, lx):
        lxa = np.array(lx, dtype=np.float64)
        lib().orc_grid_set_lx(self.h, _dp(lxa))

    @property
    def volume(self):
        return lib().orc_grid_volume(self.h)

    def integral(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return lib().orc_integral(self.h, _dp(a))

    def inner_product(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        return lib().orc_inner_product(self.h, _dp(a), _dp(b))

    def rfft(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        out = np.empty(self.Mc, dtype=np.complex128)
        lib().orc_fft_forward(self.h, _dp(a), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def irfft(self, c):
        c = np.ascontiguousarray(c, dtype=np.complex128).reshape(-1)
        out = np.empty(self.M, dtype=np.float64)
        lib().orc_fft_backward(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _dp(out))
        return out

    def boltz_bond(self, bond_length_sq, ds):
        out = np.empty(self.Mc, dtype=np.float64)
        lib().orc_boltz_bond(self.h, bond_length_sq, ds, _dp(out))
        return out

    def fourier_basis(self):
        bx, by, bz = (np.empty(self.Mc, dtype=np.float64) for _ in range(3))
        lib().orc_fourier_basis(self.h, _dp(bx), _dp(by), _dp(bz))
        return bx, by, bz


class Solver:
    """Propagators, Q, concentrations and stress for a plan on a Grid."""

    def __init__(self, grid, plan, ds, segment_lengths, solvents=(), realspace_bc=None):
        """realspace_bc: None for pseudo-spectral steps, else the boundary conditions
        (2 per axis) of real-space Crank-Nicolson steps (continuous chains)."""
        self.grid = grid
        self.plan = plan
        self.ds = ds
        self.monomers = sorted(segment_lengths)
        mon = {m: i for i, m in enumerate(self.monomers)}
        props = plan["propagators"]
        self.key_index = {p["key"]: i for i, p in enumerate(props)}
        qinit_names = []
        kind, qidx, dep_off, n_deps, dprop, dn, drep, je_off, n_je, je = ([] for _ in range(10))
        for p in props:
            key = p["key"]
            dep_off.append(len(dprop))
            n_deps.append(len(p["deps"]))
            for dk, n, r in p["deps"]:
                dprop.append(self.key_index[dk])
                dn.append(n)
                drep.append(r)
            je_off.append(len(je))
            n_je.append(len(p["junction_ends"]))
            je.extend(p["junction_ends"])
            if not p["deps"]:
                if key[0] == "{":
                    kind.append(1)
                    nm = q_input_name(key)
                    if nm not in qinit_names:
                        qinit_names.append(nm)
                    qidx.append(qinit_names.index(nm))
                else:
                    kind.append(0)
                    qidx.append(-1)
            else:
                kind.append(3 if key[0] == "[" else 2)
                qidx.append(-1)
        self.qinit_names = qinit_names
        blocks = plan["blocks"]
        self.arrays = {}

        def ia(name, v):
            a = np.array(v if v else [0], dtype=np.int32)
            self.arrays[name] = a
            return a.ctypes.data

        def da(name, v):
            a = np.array(v if v else [0.0], dtype=np.float64)
            self.arrays[name] = a
            return a.ctypes.data

        st = _Plan()
        st.model = 0 if plan["model"] == "continuous" else 1
        st.ds = ds
        st.n_monomers = len(self.monomers)
        st.segment_length = da("seg", [segment_lengths[m] for m in self.monomers])
        st.n_props = len(props)
        st.prop_monomer = ia("pm", [mon[p["monomer"]] for p in props])
        st.prop_max_n = ia("pmax", [p["max_n_segment"] for p in props])
        st.prop_kind = ia("pk", kind)
        st.prop_qinit = ia("pq", qidx)
        st.prop_dep_off = ia("pdo", dep_off)
        st.prop_n_deps = ia("pnd", n_deps)
        st.dep_prop = ia("dp", dprop)
        st.dep_n = ia("dn", dn)
        st.dep_rep = ia("dr", drep)
        st.prop_je_off = ia("pjo", je_off)
        st.prop_n_je = ia("pnj", n_je)
        st.je = ia("je", je)
        st.n_blocks = len(blocks)
        st.blk_polymer = ia("bp", [b["polymer"] for b in blocks])
        st.blk_left = ia("bl", [self.key_index[b["left"]] for b in blocks])
        st.blk_right = ia("br", [self.key_index[b["right"]] for b in blocks])
        st.blk_n_left = ia("bnl", [b["n_left"] for b in blocks])
        st.blk_n_right = ia("bnr", [b["n_right"] for b in blocks])
        st.blk_rep = ia("brep", [b["n_repeated"] for b in blocks])
        st.blk_monomer = ia("bm", [mon[b["monomer"]] for b in blocks])
        st.blk_n_vu = ia("bnvu", [len(b["v_u"]) for b in blocks])
        pols = plan["polymers"]
        st.n_polymers = len(pols)
        st.pol_vf = da("pvf", [p["volume_fraction"] for p in pols])
        st.pol_alpha = da("pal", [p["alpha"] for p in pols])
        st.n_solvents = len(solvents)
        st.sol_vf = da("svf", [s[0] for s in solvents])
        st.sol_monomer = ia("smon", [mon[s[1]] for s in solvents])
        self.solvents = list(solvents)
        self._st = st
        self.h = lib().orc_solver_create(grid.h, ctypes.byref(st))
        if realspace_bc is not None:
            if lib().orc_solver_set_realspace(self.h, _bc_array(realspace_bc, grid.dim)) != 0:
                raise RuntimeError("Real-space method only support 'continuous' chain model.")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_destroy(self.h)
            self.h = None

    def compute_propagators(self, w, q_init=None):
        M = self.grid.M
        self._w = [np.ascontiguousarray(w[m], dtype=np.float64).reshape(M) for m in self.monomers]
        wp = (ctypes.POINTER(ctypes.c_double) * len(self._w))(*[_dp(a) for a in self._w])
        qp = None
        if self.qinit_names:
            self._q = [np.ascontiguousarray(q_init[n], dtype=np.float64).reshape(M) for n in self.qinit_names]
            qp = (ctypes.POINTER(ctypes.c_double) * len(self._q))(*[_dp(a) for a in self._q])
        rc = lib().orc_compute_propagators(self.h, wp, qp)
        if rc != 0:
            raise RuntimeError("oracle: missing q_init")

    def compute_concentrations(self):
        lib().orc_compute_concentrations(self.h)

    def compute_statistics(self, w, q_init=None):
        self.compute_propagators(w, q_init)
        self.compute_concentrations()

    def compute_stress(self):
        lib().orc_compute_stress(self.h)

    def get_total_partition(self, p):
        return lib().orc_partition(self.h, p)

    def get_solvent_partition(self, s):
        return lib().orc_solvent_partition(self.h, s)

    def block_phi(self, b):
        return np.ctypeslib.as_array(lib().orc_block_phi(self.h, b), shape=(self.grid.M,)).copy()

    def get_total_concentration(self, monomer, polymer=None):
        phi = np.zeros(self.grid.M)
        for b, blk in enumerate(self.plan["blocks"]):
            if polymer is not None and blk["polymer"] != polymer:
                continue
            if blk["monomer"] == monomer and blk["n_right"] != 0:
                phi += self.block_phi(b)
        if polymer is None:
            for i, (vf, m) in enumerate(self.solvents):
                if m == monomer:
                    phi += self.get_solvent_concentration(i)
        return phi

    def get_solvent_concentration(self, s):
        return np.ctypeslib.as_array(lib().orc_solvent_phi(self.h, s), shape=(self.grid.M,)).copy()

    def dq_dl(self, p):
        return [lib().orc_dq_dl(self.h, p, d) for d in range(self.grid.dim)]

    def get_stress(self):
        """PropagatorComputation.cpp:35-52"""
        out = np.zeros(self.grid.dim)
        for p, pol in enumerate(self.plan["polymers"]):
            out += np.array(self.dq_dl(p)) * pol["volume_fraction"] / pol["alpha"] / self.get_total_partition(p)
        return out

    def propagator(self, key, n, half=False):
        ptr = lib().orc_propagator(self.h, self.key_index[key], n, 1 if half else 0)
        if not ptr:
            raise KeyError((key, n, half))
        return np.ctypeslib.as_array(ptr, shape=(self.grid.M,)).copy()

    def get_chain_propagator(self, p, v, u, n):
        for blk in self.plan["polymers"][p]["blocks"]:
            if (blk["v"], blk["u"]) == (v, u):
                return self.propagator(blk["key_vu"], n)
            if (blk["u"], blk["v"]) == (v, u):
                return self.propagator(blk["key_uv"], n)
        raise KeyError((p, v, u))

    def block_n_segment(self, p, v, u):
        for blk in self.plan["polymers"][p]["blocks"]:
            if {blk["v"], blk["u"]} == {v, u}:
                return blk["n_segment"]
        raise KeyError((p, v, u))


class AndersonMixing:
    def __init__(self, n_var, max_hist, start_error, mix_min, mix_init):
        self.n_var = n_var
        self.h = lib().orc_am_create(n_var, max_hist, start_error, mix_min, mix_init)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_am_destroy(self.h)
            self.h = None

    def reset_count(self):
        lib().orc_am_reset(self.h)

    def calculate_new_fields(self, w, d, old_err, err):
        w = np.ascontiguousarray(w, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        out = np.empty(self.n_var)
        lib().orc_am_new_fields(self.h, _dp(out), _dp(w), _dp(d), old_err, err)
        return out
