"""ctypes binding of libfts_hip.so (include/fts_hip.h).

This is the plain-C entry to the HIP path for callers that hold device pointers
(torch tensors, their own hipMalloc) or that build the propagator plan
themselves. The reference-shaped Python API (PlatformSelector, factories,
PropagatorComputation, ...) is the pybind11 module langevin_fts_amd.langevinfts.

The library must be built in-tree (python -m langevin_fts_amd.build); loading
fails loudly otherwise.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FTS_CONTINUOUS, FTS_DISCRETE = 0, 1
FTS_FP64, FTS_FP32 = 0, 1
FTS_LEAF, FTS_LEAF_QINIT, FTS_JUNCTION, FTS_AGGREGATED = 0, 1, 2, 3
FTS_SRC_BLOCK, FTS_SRC_SOLVENT = 0, 1

_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)


class FtsPlan(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("ds", ctypes.c_double), ("n_monomers", ctypes.c_int),
                ("segment_length", _DP), ("n_props", ctypes.c_int)] + \
        [(n, _IP) for n in ["prop_monomer", "prop_max_n", "prop_kind", "prop_qinit", "prop_dep_off",
                            "prop_n_deps", "dep_prop", "dep_n", "dep_rep", "prop_je_off", "prop_n_je", "je"]] + \
        [("n_blocks", ctypes.c_int)] + \
        [(n, _IP) for n in ["blk_polymer", "blk_left", "blk_right", "blk_n_left", "blk_n_right", "blk_rep",
                            "blk_monomer", "blk_n_vu"]] + \
        [("n_polymers", ctypes.c_int), ("pol_vf", _DP), ("pol_alpha", _DP),
         ("n_solvents", ctypes.c_int), ("sol_vf", _DP), ("sol_monomer", _IP),
         ("n_qinit", ctypes.c_int), ("n_slices", ctypes.c_int), ("slice_off", _IP),
         ("job_prop", _IP), ("job_from", _IP), ("job_to", _IP)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libfts_hip.so")
        if not os.path.exists(path):
            raise RuntimeError("libfts_hip.so is not built; run `python -m langevin_fts_amd.build`")
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        VPP = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "fts_last_error": (ctypes.c_char_p, []),
            "fts_abi_version": (i, []),
            "fts_context_create": (i, [i, VPP]),
            "fts_context_destroy": (i, [vp]),
            "fts_context_synchronize": (i, [vp]),
            "fts_context_stream": (i, [vp, VPP]),
            "fts_context_info": (i, [vp, ctypes.c_char_p, i]),
            "fts_context_set_kernel_timing": (i, [vp, i]),
            "fts_grid_create": (i, [vp, i, _IP, _DP, _DP, VPP]),
            "fts_grid_destroy": (i, [vp]),
            "fts_grid_set_lx": (i, [vp, _DP]),
            "fts_grid_volume": (i, [vp, _DP]),
            "fts_grid_integral": (i, [vp, _DP, i, _DP]),
            "fts_grid_inner_product": (i, [vp, _DP, _DP, i, _DP]),
            "fts_grid_fft_forward": (i, [vp, i, vp, vp, i]),
            "fts_grid_fft_backward": (i, [vp, i, vp, vp, i]),
            "fts_grid_fft_backward_noalloc": (i, [vp, i, vp, vp, i]),
            "fts_saddle_set_graphs": (i, [vp, i]),
            "fts_grid_get_fft_schedule": (i, [vp, i, _IP, _IP, ctypes.c_char_p, i]),
            "fts_grid_set_fft_schedule": (i, [vp, i, i, i]),
            "fts_solver_create": (i, [vp, ctypes.POINTER(FtsPlan), i, VPP]),
            "fts_solver_create_realspace": (i, [vp, ctypes.POINTER(FtsPlan), _IP, VPP]),
            "fts_solver_create_reduce_memory": (i, [vp, ctypes.POINTER(FtsPlan), i, _IP, VPP]),
            "fts_solver_get_memory_usage": (i, [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
            "fts_am_create_reduce_memory": (i, [vp, i, i, d, d, d, VPP]),
            "fts_solver_destroy": (i, [vp]),
            "fts_solver_update_laplacian": (i, [vp]),
            "fts_solver_compute_propagators": (i, [vp, ctypes.POINTER(_DP), ctypes.POINTER(_DP), i]),
            "fts_solver_compute_concentrations": (i, [vp]),
            "fts_solver_set_graphs": (i, [vp, i]),
            "fts_solver_compute_stress": (i, [vp]),
            "fts_solver_get_partition": (i, [vp, i, _DP]),
            "fts_solver_get_solvent_partition": (i, [vp, i, _DP]),
            "fts_solver_get_dq_dl": (i, [vp, i, _DP]),
            "fts_solver_sum_concentrations": (i, [vp, i, _IP, _IP, _DP, _DP, i]),
            "fts_solver_get_propagator": (i, [vp, i, i, i, _DP, i]),
            "fts_am_create": (i, [vp, i, i, d, d, d, VPP]),
            "fts_am_destroy": (i, [vp]),
            "fts_am_reset": (i, [vp]),
            "fts_am_calculate_new_fields": (i, [vp, _DP, _DP, _DP, d, d, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().fts_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def as_dp(addr):
    """A device address (int) as double*."""
    return ctypes.cast(ctypes.c_void_p(addr), _DP)


_HIPRT = None


def hip_runtime():
    """The HIP runtime libfts_hip.so is linked against (already loaded with it).

    Device buffers handed to the library must come from this runtime; a second
    HIP runtime in the process (such as the one bundled with a framework wheel)
    does not share devices or allocations with it."""
    global _HIPRT
    if _HIPRT is None:
        lib()
        R = ctypes.CDLL("libamdhip64.so.7")
        R.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        R.hipFree.argtypes = [ctypes.c_void_p]
        R.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        R.hipDeviceSynchronize.argtypes = []
        _HIPRT = R
    return _HIPRT


class DeviceArray:
    """A device allocation of the library's HIP runtime, with host copies."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        if hip_runtime().hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if hip_runtime().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy host to device failed")

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if hip_runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy device to host failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _HIPRT is not None:
            _HIPRT.hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().fts_context_create(device, ctypes.byref(h)))
        self.h = h

    def synchronize(self):
        check(lib().fts_context_synchronize(self.h))

    def set_kernel_timing(self, sample_every):
        """time one launch in sample_every of each kernel family with HIP events (0: off)"""
        check(lib().fts_context_set_kernel_timing(self.h, sample_every))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            lib().fts_context_destroy(self.h)
            self.h = None


class Grid:
    def __init__(self, ctx, nx, lx, mask=None):
        self.ctx = ctx
        self.nx = list(nx)
        self.dim = len(nx)
        self.M = int(np.prod(nx))
        self.Mc = int(np.prod(nx[:-1])) * (nx[-1] // 2 + 1)
        nxa = np.array(nx, dtype=np.int32)
        lxa = np.array(lx, dtype=np.float64)
        self._mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().fts_grid_create(ctx.h, self.dim, _ip(nxa), _dp(lxa),
                                    None if self._mask is None else _dp(self._mask), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            lib().fts_grid_destroy(self.h)
            self.h = None

    def set_lx(self, lx):
        check(lib().fts_grid_set_lx(self.h, _dp(np.array(lx, dtype=np.float64))))

    @property
    def volume(self):
        v = ctypes.c_double()
        check(lib().fts_grid_volume(self.h, ctypes.byref(v)))
        return v.value

    def integral(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        v = ctypes.c_double()
        check(lib().fts_grid_integral(self.h, _dp(a), 0, ctypes.byref(v)))
        return v.value

    def inner_product(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        v = ctypes.c_double()
        check(lib().fts_grid_inner_product(self.h, _dp(a), _dp(b), 0, ctypes.byref(v)))
        return v.value

    def fft_forward(self, d_in, d_out, batch=1, precision=FTS_FP64):
        check(lib().fts_grid_fft_forward(self.h, precision, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), batch))

    def fft_backward(self, d_in, d_out, batch=1, precision=FTS_FP64):
        check(lib().fts_grid_fft_backward(self.h, precision, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), batch))

    def fft_schedule(self, precision=FTS_FP64):
        """(slab workgroup threads, pencil kz columns per wavefront, note) of the
        wave-local FFT kernels; (0, 0, ...) on grids without them."""
        bd, nl = ctypes.c_int(), ctypes.c_int()
        buf = ctypes.create_string_buffer(512)
        check(lib().fts_grid_get_fft_schedule(self.h, precision, ctypes.byref(bd), ctypes.byref(nl), buf, 512))
        return bd.value, nl.value, buf.value.decode()

    def set_fft_schedule(self, slab_threads, pencil_columns, precision=FTS_FP64):
        check(lib().fts_grid_set_fft_schedule(self.h, precision, slab_threads, pencil_columns))


def q_input_name(key):
    return key[1:key.rindex("}")]


class Plan:
    """fts_plan from the analyzer/scheduler description (the dict form used by the
    tests' golden files: propagators, blocks, polymers and a schedule)."""

    def __init__(self, plan, ds, segment_lengths, schedule, solvents=()):
        self.monomers = sorted(segment_lengths)
        mon = {m: i for i, m in enumerate(self.monomers)}
        props = plan["propagators"]
        self.key_index = {p["key"]: i for i, p in enumerate(props)}
        self.qinit_names = []
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
                    name = q_input_name(key)
                    if name not in self.qinit_names:
                        self.qinit_names.append(name)
                    kind.append(FTS_LEAF_QINIT)
                    qidx.append(self.qinit_names.index(name))
                else:
                    kind.append(FTS_LEAF)
                    qidx.append(-1)
            else:
                kind.append(FTS_AGGREGATED if key[0] == "[" else FTS_JUNCTION)
                qidx.append(-1)
        blocks = plan["blocks"]
        slice_off, jp, jf, jt = [0], [], [], []
        for sl in schedule:
            for key, f, t in sl:
                jp.append(self.key_index[key])
                jf.append(f)
                jt.append(t)
            slice_off.append(len(jp))
        self._keep = []

        def ia(v):
            a = np.array(v if len(v) else [0], dtype=np.int32)
            self._keep.append(a)
            return _ip(a)

        def da(v):
            a = np.array(v if len(v) else [0.0], dtype=np.float64)
            self._keep.append(a)
            return _dp(a)

        s = FtsPlan()
        s.model = FTS_CONTINUOUS if plan["model"] == "continuous" else FTS_DISCRETE
        s.ds = ds
        s.n_monomers = len(self.monomers)
        s.segment_length = da([segment_lengths[m] for m in self.monomers])
        s.n_props = len(props)
        s.prop_monomer = ia([mon[p["monomer"]] for p in props])
        s.prop_max_n = ia([p["max_n_segment"] for p in props])
        s.prop_kind = ia(kind)
        s.prop_qinit = ia(qidx)
        s.prop_dep_off = ia(dep_off)
        s.prop_n_deps = ia(n_deps)
        s.dep_prop = ia(dprop)
        s.dep_n = ia(dn)
        s.dep_rep = ia(drep)
        s.prop_je_off = ia(je_off)
        s.prop_n_je = ia(n_je)
        s.je = ia(je)
        s.n_blocks = len(blocks)
        s.blk_polymer = ia([b["polymer"] for b in blocks])
        s.blk_left = ia([self.key_index[b["left"]] for b in blocks])
        s.blk_right = ia([self.key_index[b["right"]] for b in blocks])
        s.blk_n_left = ia([b["n_left"] for b in blocks])
        s.blk_n_right = ia([b["n_right"] for b in blocks])
        s.blk_rep = ia([b["n_repeated"] for b in blocks])
        s.blk_monomer = ia([mon[b["monomer"]] for b in blocks])
        s.blk_n_vu = ia([len(b["v_u"]) for b in blocks])
        pols = plan["polymers"]
        s.n_polymers = len(pols)
        s.pol_vf = da([p["volume_fraction"] for p in pols])
        s.pol_alpha = da([p["alpha"] for p in pols])
        s.n_solvents = len(solvents)
        s.sol_vf = da([v for v, _ in solvents])
        s.sol_monomer = ia([mon[m] for _, m in solvents])
        s.n_qinit = len(self.qinit_names)
        s.n_slices = len(schedule)
        s.slice_off = ia(slice_off)
        s.job_prop = ia(jp)
        s.job_from = ia(jf)
        s.job_to = ia(jt)
        self.struct = s
        self.plan = plan
        self.solvents = list(solvents)


BC_CODES = {"periodic": 0, "reflecting": 1, "absorbing": 2}


class Solver:
    def __init__(self, grid, plan, precision=FTS_FP64, realspace_bc=None):
        """realspace_bc: None for pseudo-spectral steps, else 2 per axis of
        "periodic", "reflecting", "absorbing" (real-space steps, FP64)."""
        self.grid = grid
        self.plan = plan
        h = ctypes.c_void_p()
        if realspace_bc is None:
            check(lib().fts_solver_create(grid.h, ctypes.byref(plan.struct), precision, ctypes.byref(h)))
        else:
            bc = np.array([BC_CODES[b] for b in realspace_bc], dtype=np.int32)
            check(lib().fts_solver_create_realspace(grid.h, ctypes.byref(plan.struct), _ip(bc), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            lib().fts_solver_destroy(self.h)
            self.h = None

    def update_laplacian(self):
        check(lib().fts_solver_update_laplacian(self.h))

    def compute_propagators(self, w, q_init=None):
        """w: {monomer: host float64 array}"""
        M = self.grid.M
        self._w = [np.ascontiguousarray(w[m], dtype=np.float64).reshape(M) for m in self.plan.monomers]
        wp = (_DP * len(self._w))(*[_dp(a) for a in self._w])
        qp = None
        if self.plan.qinit_names:
            self._q = [np.ascontiguousarray(q_init[n], dtype=np.float64).reshape(M) for n in self.plan.qinit_names]
            qp = (_DP * len(self._q))(*[_dp(a) for a in self._q])
        check(lib().fts_solver_compute_propagators(self.h, wp, qp, 0))

    def compute_propagators_device(self, w_ptrs, q_ptrs=()):
        """w_ptrs: device addresses (int) of FP64 fields in monomer order"""
        wp = (_DP * len(w_ptrs))(*[as_dp(a) for a in w_ptrs])
        qp = (_DP * len(q_ptrs))(*[as_dp(a) for a in q_ptrs]) if q_ptrs else None
        check(lib().fts_solver_compute_propagators(self.h, wp, qp, 1))

    def compute_concentrations(self):
        check(lib().fts_solver_compute_concentrations(self.h))

    def set_graphs(self, enable=True):
        """replay device-input evaluations from captured HIP graphs"""
        check(lib().fts_solver_set_graphs(self.h, 1 if enable else 0))

    def compute_statistics(self, w, q_init=None):
        self.compute_propagators(w, q_init)
        self.compute_concentrations()

    def compute_stress(self):
        check(lib().fts_solver_compute_stress(self.h))

    def get_total_partition(self, p):
        v = ctypes.c_double()
        check(lib().fts_solver_get_partition(self.h, p, ctypes.byref(v)))
        return v.value

    def get_solvent_partition(self, s):
        v = ctypes.c_double()
        check(lib().fts_solver_get_solvent_partition(self.h, s, ctypes.byref(v)))
        return v.value

    def dq_dl(self, p):
        out = np.zeros(3)
        check(lib().fts_solver_get_dq_dl(self.h, p, _dp(out)))
        return out[:self.grid.dim]

    def get_stress(self):
        out = np.zeros(self.grid.dim)
        for p, pol in enumerate(self.plan.plan["polymers"]):
            out += self.dq_dl(p) * pol["volume_fraction"] / pol["alpha"] / self.get_total_partition(p)
        return out

    def sum_concentrations(self, kinds, idx, scale=None, out=None):
        kinds = np.array(kinds, dtype=np.int32)
        idx = np.array(idx, dtype=np.int32)
        sc = None if scale is None else np.array(scale, dtype=np.float64)
        res = np.empty(self.grid.M) if out is None else out
        check(lib().fts_solver_sum_concentrations(self.h, len(kinds), _ip(kinds), _ip(idx),
                                                  None if sc is None else _dp(sc), _dp(res), 0))
        return res

    def get_total_concentration(self, monomer, polymer=None):
        kinds, idx = [], []
        for b, blk in enumerate(self.plan.plan["blocks"]):
            if polymer is not None and blk["polymer"] != polymer:
                continue
            if blk["monomer"] == monomer and blk["n_right"] != 0:
                kinds.append(FTS_SRC_BLOCK)
                idx.append(b)
        if polymer is None:
            for s, (vf, m) in enumerate(self.plan.solvents):
                if m == monomer:
                    kinds.append(FTS_SRC_SOLVENT)
                    idx.append(s)
        return self.sum_concentrations(kinds, idx)

    def get_solvent_concentration(self, s):
        return self.sum_concentrations([FTS_SRC_SOLVENT], [s])

    def propagator(self, key, n, half=False):
        out = np.empty(self.grid.M)
        check(lib().fts_solver_get_propagator(self.h, self.plan.key_index[key], n, 1 if half else 0, _dp(out), 0))
        return out

    def get_chain_propagator(self, p, v, u, n):
        for blk in self.plan.plan["polymers"][p]["blocks"]:
            if (blk["v"], blk["u"]) == (v, u):
                return self.propagator(blk["key_vu"], n)
            if (blk["u"], blk["v"]) == (v, u):
                return self.propagator(blk["key_uv"], n)
        raise KeyError((p, v, u))


class AndersonMixing:
    def __init__(self, ctx, n_var, max_hist, start_error, mix_min, mix_init):
        self.n_var = n_var
        h = ctypes.c_void_p()
        check(lib().fts_am_create(ctx.h, n_var, max_hist, start_error, mix_min, mix_init, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            lib().fts_am_destroy(self.h)
            self.h = None

    def reset_count(self):
        check(lib().fts_am_reset(self.h))

    def calculate_new_fields(self, w, d, old_err, err):
        w = np.ascontiguousarray(w, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        out = np.empty(self.n_var)
        check(lib().fts_am_calculate_new_fields(self.h, _dp(out), _dp(w), _dp(d), old_err, err, 0))
        return out
