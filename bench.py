"""Langevin FTS throughput on one MI355X, with the propagator kernels' roofline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lfts64|lfts32|replicas48|realspace3d|reduce128|c1|c2|c2d|c3|c4]

Workload "lfts64" (default): AB diblock melt, discrete chain N=90, f=0.5, chiN=17,
64^3 periodic grid, L=6.0 R0, nbar=1e4, Langevin step dt=0.8 with PCG64 noise,
saddle point to 1e-4 by Anderson mixing (history 20, start 0.5, mix 0.01/0.01),
everything FP64. A "step" is one Langevin step: the Leimkuhler-Matthews update of
w- and the saddle-point iteration of w+ (each iteration = all propagators,
concentrations and one Anderson mixing), as examples/fts/lfts.py:687-731 does it.
Fields start as N(0,1) from a fixed seed; the initial saddle point and W warm-up
steps run untimed. Then, with the same tolerance:

  training     --guess-train-steps steps (default 1000) that train the CNN guess on
               the run's own saddle points (timed separately, reported, not in value);
  plain arm    K timed steps, each saddle point starting from the previous w+, all
               iterations FP64 (the reference's method);
  learned FP64 K steps whose saddle points start from the CNN's prediction, FP64;
  plain mixed  K steps from the reference's start with the mixed-precision saddle
               point: FP32 propagators while the error is >= tol, FP64 after; only
               FP64 evaluations can declare convergence, so every saddle point
               meets the same FP64 tolerance;
  learned      K steps with both (skipped with --precision fp64).

`value` is the learned-guess arm with FP64 saddle iterations (the reference's
precision); config carries every arm's rate and iterations per step, the ratio
to the plain arm, the mixed-precision arm with its largest (FP64) saddle error,
the rate of a 10,000-step run that trains on its first steps, and the agreement
of the structure functions S_ij(k) (lfts.py:753-766) of the plain and learned
FP64 arms run from one state with one noise sequence (untimed, --sk-steps).

Printed: one JSON line. `roofline` is the slab pass (the dominant kernel:
inverse y/z FFT, exp(-w ds) factor, q store, forward z/y FFT), timed during the
timed steps with fence-free HIP events around one launch in 16 on the library's
stream; its algorithmic bytes are the spectra, fields and w each launch must move
once. `traffic` is the slab launch's HBM bytes from two `rocprofv3 --pmc` child
runs (FETCH_SIZE, WRITE_SIZE) made before this process touches the GPU, each
scaled by what the same counter reads for a plain device copy of 256 MiB.
`cpu_baseline` times saddle iterations of the same 64^3 workload on the CPU in
the reference CPU path's structure (oracle/cpu_pocketfft.py: a slice's propagators
as parallel jobs, each with sequential pocketfft FFTs), over the whole contour, and
scales by the GPU run's iterations per step.
"""
import argparse
import contextlib
import csv
import glob
import itertools
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

PEAK_HBM_GBPS = 8000.0
SAMPLE = 16   # one launch in SAMPLE of each kernel family carries timing events

WORKLOADS = {
    "lfts64": dict(nx=[64, 64, 64], lx=[6.0, 6.0, 6.0], model="discrete", ds=1 / 90, f=0.5, chi_n=17.0,
                   nbar=1e4, dt=0.8, tol=1e-4, am=(20, 5e-1, 0.01, 0.01), max_iter=200),
    "lfts32": dict(nx=[32, 32, 32], lx=[4.0, 4.0, 4.0], model="discrete", ds=1 / 90, f=0.5, chi_n=17.0,
                   nbar=1e4, dt=0.8, tol=1e-4, am=(20, 5e-1, 0.01, 0.01), max_iter=200),
    # batching form: 32 independent replicas advanced together
    "replicas48": dict(nx=[48, 48, 48], lx=[4.0, 4.0, 4.0], model="discrete", ds=1 / 64, f=0.5, chi_n=17.0,
                       nbar=1e4, dt=0.8, tol=1e-4, am=(20, 5e-1, 0.01, 0.01), max_iter=200, replicas=32),
    # reduce-memory form: propagators in pinned host memory (the 128^3 discrete N=90 precision-form shape)
    "reduce128": dict(nx=[128, 128, 128], lx=[6.0, 6.0, 6.0], model="discrete", ds=1 / 90, f=0.4, reduce=True),
    # propagator evaluations of the measurement plan's other configurations
    "c1": dict(kind="eval", nx=[32, 32, 32], lx=[4.0, 4.0, 4.0], model="continuous", ds=0.01, sigma=1.0,
               bonds={"A": 1.0, "B": 1.0}, polymers=[[1.0, [["A", 0.5, 0, 1], ["B", 0.5, 1, 2]]]],
               label="C1: AB diblock, continuous, f=0.5, ds=0.01, 32^3, L=4"),
    "c3": dict(kind="eval", nx=[128, 128, 128], lx=[6.0, 6.0, 6.0], model="continuous", ds=0.01, sigma=2.0,
               bonds={"A": 1.0, "B": 1.0}, polymers=[[1.0, [["A", 0.4, 0, 1], ["B", 0.6, 1, 2]]]],
               precisions=["hip-fp32", "hip"],
               label="C3: AB diblock, continuous, f=0.4, ds=0.01, 128^3, L=6, w ~ N(0, 2^2)"),
    "c4": dict(kind="eval", nx=[64, 64, 64], lx=[4.0, 4.0, 4.0], model="continuous", ds=0.01, sigma=1.0,
               bonds={"A": 1.0, "B": 1.0, "C": 1.2},
               polymers=[[0.8, [["A", 0.3, 0, 1], ["B", 0.3, 0, 2], ["C", 0.4, 0, 3]]], [0.2, [["A", 0.5, 0, 1]]]],
               label="C4: ABC star (0.3/0.3/0.4, b=1/1/1.2, phi=0.8) + A homopolymer (0.5, phi=0.2), continuous, "
                     "ds=0.01, 64^3, L=4"),
    # SCFT iterations (propagators, concentrations, stress, Anderson mixing on [w, L], box relaxation)
    "c2": dict(kind="scft", nx=[64, 64, 64], lx=[3.3, 3.3, 3.3], model="continuous", ds=0.01, f=0.36, chi_n=20.0,
               label="C2: SCFT of an AB diblock, continuous, f=0.36, chiN=20, ds=0.01, 64^3, L=3.3, box relaxation"),
    "c2d": dict(kind="scft", nx=[64, 64, 64], lx=[3.3, 3.3, 3.3], model="discrete", ds=0.01, f=0.36, chi_n=20.0,
                label="C2: SCFT of an AB diblock, discrete N=100, f=0.36, chiN=20, 64^3, L=3.3, box relaxation"),
    # real-space Crank-Nicolson solver: the reference's 3D nanoparticle example (devel/RealSpace3D.py)
    "realspace3d": dict(nx=[120, 100, 140], lx=[4.0, 5.0, 6.0], model="continuous", ds=0.01, radius=1.0,
                        bc=["absorbing", "absorbing", "reflecting", "reflecting", "reflecting", "reflecting"],
                        realspace=True),
}


@contextlib.contextmanager
def stdout_to_stderr():
    """Send everything written to fd 1 (Python and the C++ layer) to fd 2."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        with contextlib.redirect_stdout(sys.stderr):
            yield
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def describe(name, w):
    n = round(1 / w["ds"])
    rep = f"{w['replicas']} replicas of " if w.get("replicas", 1) > 1 else ""
    return (f"{rep}L-FTS AB diblock, {w['model']} N={n}, f={w['f']}, chiN={w['chi_n']}, "
            f"{'x'.join(map(str, w['nx']))} grid, L={w['lx'][0]}, nbar={w['nbar']:g}, dt={w['dt']}, "
            f"saddle tol {w['tol']:g}")


def lfts_params(w, rec_dir):
    return {
        "nx": w["nx"], "lx": w["lx"], "chain_model": w["model"], "ds": w["ds"],
        "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": w["chi_n"]},
        "distinct_polymers": [{"volume_fraction": 1.0, "blocks": [{"type": "A", "length": w["f"]},
                                                                 {"type": "B", "length": 1 - w["f"]}]}],
        "langevin": {"max_step": 0, "dt": w["dt"], "nbar": w["nbar"]},
        "recording": {"dir": rec_dir, "recording_period": 10 ** 9, "sf_computing_period": 10 ** 9,
                      "sf_recording_period": 10 ** 9},
        "saddle": {"max_iter": w["max_iter"], "tolerance": w["tol"]},
        "am": {"max_hist": w["am"][0], "start_error": w["am"][1], "mix_min": w["am"][2], "mix_init": w["am"][3]},
        "verbose_level": 0,
    }


def langevin_steps(sim, n, stats, learned=False):
    """n Langevin steps on the device state (lfts.py:687-731); with learned=True the
    driver's learned-guess step (training until the network is ready, then prediction)."""
    dev = sim.saddle_device
    tol = sim.saddle["tolerance"]
    for _ in range(n):
        noise = sim.next_noise()   # drawn ahead while the previous saddle point ran
        if learned:
            it, err, ok = sim.learned_step(noise)
        else:
            dev.backup()
            dev.langevin_update(noise, sim.langevin["dt"])
            it, err, _ = sim._find_saddle_point_device()
            ok = not (np.isnan(err) or err >= tol)
            if not ok:
                dev.restore()
        stats["iterations"] += it
        stats["failures"] += 0 if ok else 1
        if ok:
            stats["max_error"] = max(stats.get("max_error", 0.0), float(err))


def full_run(args, sim, precision, w):
    """--full-run N: one run of N Langevin steps timed end to end, as a user's run would go:
    the first --guess-train-steps steps train the guess (mixed-precision saddle points),
    the rest use it with FP64 saddle iterations. Progress goes to stderr every 250 steps."""
    n, n_train = args.full_run, args.guess_train_steps
    stats_train = {"iterations": 0, "failures": 0}
    stats_run = {"iterations": 0, "failures": 0}
    sim.saddle_device.field_moments()
    t0 = time.perf_counter()
    precision(True)
    for k in range(0, n_train, 250):
        langevin_steps(sim, min(250, n_train - k), stats_train, learned=True)
        print(f"full run: {k + min(250, n_train - k)} training steps, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    precision(False)
    t1 = time.perf_counter()
    for k in range(0, n - n_train, 250):
        langevin_steps(sim, min(250, n - n_train - k), stats_run, learned=True)
        print(f"full run: {n_train + k + min(250, n - n_train - k)} steps, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    sim.saddle_device.field_moments()
    t2 = time.perf_counter()
    out = {
        "metric": f"L-FTS steps/s over one whole {n}-step run with the learned guess (training included)",
        "value": n / (t2 - t0), "unit": "steps/s", "n_gpus": 1, "steps": n, "warmup": args.warmup,
        "ms_per_step": 1e3 * (t2 - t0) / n, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (training steps: f32 propagators until err < tol, then f64)",
        "data": "synthetic: w_A, w_B ~ N(0,1) from PCG64(seed); Langevin noise from numpy PCG64",
        "config": {"workload": describe(args.workload, w), "train_steps": n_train, "train_s": t1 - t0,
                   "train_saddle_iterations_per_step": stats_train["iterations"] / max(n_train, 1),
                   "learned_steps_per_s": (n - n_train) / (t2 - t1),
                   "learned_saddle_iterations_per_step": stats_run["iterations"] / max(n - n_train, 1),
                   "saddle_failures": stats_train["failures"] + stats_run["failures"],
                   "train_loss_last50": float(np.mean(sim.guess_losses[-50:])) if sim.guess_losses else None},
    }
    return out


def baseline_plan(w, ds=None):
    """The propagator plan of the workload's molecules from the product's analyzer
    (host code only), in the oracle's plan form."""
    from langevin_fts_amd import langevinfts as lf

    bonds = w.get("bonds", {"A": 1.0, "B": 1.0})
    polymers = w.get("polymers", [[1.0, [["A", w.get("f", 0.5), 0, 1], ["B", 1 - w.get("f", 0.5), 1, 2]]]])
    factory = lf.PlatformSelector.create_factory("hip", False)
    mol = factory.create_molecules_information(w["model"], ds or w["ds"], bonds)
    for vf, blocks in polymers:
        mol.add_polymer(vf, blocks)
    an = factory.create_propagator_analyzer(mol, True)
    return plan_dict(mol, an, w["model"]), bonds


def baseline_solver(w):
    """The CPU baseline in the reference CPU path's structure: the propagators of a
    slice as parallel jobs, each with its own sequential FFTs (pocketfft)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import cpu_pocketfft as P

    plan, bonds = baseline_plan(w)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    solver = P.PocketSolver(w["nx"], w["lx"], plan, w["ds"], bonds, threads=threads)
    jobs = max(len(level) for level in solver.levels)
    return solver, bonds, min(threads, max(jobs, len(plan["blocks"])))


def baseline_note(solver, cores):
    return (f"FP64, reference CPU structure: the propagators of a schedule slice as parallel jobs (at most "
            f"{max(len(lv) for lv in solver.levels)} at once here), each job's FFTs sequential (scipy.fft pocketfft, "
            f"1 worker), block concentrations in parallel; thread pool of {solver.threads} on the host, "
            f"{cores} busy at most; {os.cpu_count()} host cores visible")


def structure_function_agreement(sim, n_steps, seed):
    """Plain and learned-guess FP64 saddle points from the same state with the same
    noise sequence: each arm accumulates the structure functions S_ij(k) of
    lfts.py:753-766 over n_steps Langevin steps. Reports the relative L2 difference of
    every S_ij(k) (k != 0) between the arms, of their spherical averages, and the
    largest field difference at the end."""
    dev = sim.saddle_device
    R = len(sim.mpt.aux_fields_real_idx)
    M = sim.cb.get_n_grid()
    w0, n0 = dev.get_fields(), dev.get_noise_prev()
    g = np.random.Generator(np.random.PCG64(seed + 1))
    noises = [g.normal(0.0, sim.langevin["sigma"], [R, M]) for _ in range(n_steps)]
    acc, fields, iters = {}, {}, {}
    for arm in ("plain", "learned"):
        dev.set_fields(np.ascontiguousarray(w0))
        dev.set_noise_prev(np.ascontiguousarray(n0))
        total, its = None, 0
        for noise in noises:
            if arm == "learned":
                it, err, ok = sim.learned_step(noise)
            else:
                dev.backup()
                dev.langevin_update(noise, sim.langevin["dt"])
                it, err, _ = sim._find_saddle_point_device()
                ok = not (np.isnan(err) or err >= sim.saddle["tolerance"])
            if not ok:
                raise RuntimeError("a saddle point of the agreement pass missed the tolerance")
            its += it
            terms = sim.structure_function_terms(dev.get_fields(), sim._phi_from_device())
            total = terms if total is None else {k: total[k] + v for k, v in terms.items()}
        acc[arm], fields[arm], iters[arm] = total, dev.get_fields(), its / n_steps
    nx = sim.cb.get_nx()
    k2 = sum(np.meshgrid(*[np.minimum(np.arange(n), n - np.arange(n)) ** 2 if d < len(nx) - 1
                           else np.arange(n // 2 + 1) ** 2 for d, n in enumerate(nx)], indexing="ij"))
    shell = np.rint(np.sqrt(k2)).astype(int).ravel()
    out = {"steps": n_steps, "saddle_iterations_per_step": iters, "pairs": {}}
    for key in acc["plain"]:
        a, b = acc["plain"][key].ravel().copy(), acc["learned"][key].ravel().copy()
        a[0] = b[0] = 0.0
        sa = np.bincount(shell, weights=a.real) / np.maximum(np.bincount(shell), 1)
        sb = np.bincount(shell, weights=b.real) / np.maximum(np.bincount(shell), 1)
        out["pairs"][key] = {"rel_l2": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                             "spherical_rel_l2": float(np.linalg.norm(sb[1:] - sa[1:]) / np.linalg.norm(sa[1:]))}
    out["max_field_difference"] = float(np.max(np.abs(fields["learned"] - fields["plain"])))
    return out


def cpu_baseline(w, iterations_per_step, budget_s):
    """Saddle iterations of the same workload on the CPU baseline (propagators,
    concentrations, residual and the oracle's Anderson mixing), timed over the
    real contour, times the GPU run's iterations per Langevin step."""
    from oracle import oracle as O

    solver, _, cores = baseline_solver(w)
    M = int(np.prod(w["nx"]))
    am = O.AndersonMixing(M, *w["am"])
    rng = np.random.default_rng(1)
    w_aux = np.stack([rng.standard_normal(M), rng.standard_normal(M)])
    n, t0 = 0, time.perf_counter()
    err = 1e20
    while True:
        solver.compute_propagators({"A": w_aux[0] + w_aux[1], "B": -w_aux[0] + w_aux[1]})
        solver.compute_concentrations()
        h = -(solver.get_total_concentration("A") + solver.get_total_concentration("B") - 1.0)
        old, err = err, np.std(h)
        w_aux[1] = am.calculate_new_fields(w_aux[1], -h, old, err)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t_iter = (time.perf_counter() - t0) / n
    return {"value": 1.0 / (t_iter * iterations_per_step), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{n} saddle iteration(s) of the same {w['nx'][0]}^3 workload, whole contour "
                      f"({t_iter:.2f} s each) x {iterations_per_step:.1f} iterations per Langevin step measured "
                      f"on the GPU; " + baseline_note(solver, cores)}


def plan_dict(mol, an, model):
    """The analyzer's propagators and blocks in the oracle's plan form."""
    props = []
    for key, e in an.get_computation_propagators().items():
        props.append({"key": key, "max_n_segment": e.max_n_segment, "monomer": e.monomer_type,
                      "deps": [list(d) for d in e.deps], "junction_ends": sorted(e.junction_ends)})
    blocks = []
    for (p, left, right), b in an.get_computation_blocks().items():
        blocks.append({"polymer": p, "left": left, "right": right, "monomer": b.monomer_type,
                       "n_left": b.n_segment_left, "n_right": b.n_segment_right, "n_repeated": b.n_repeated,
                       "v_u": [list(x) for x in b.v_u]})
    pols = [{"alpha": mol.get_polymer(p).get_alpha(), "volume_fraction": mol.get_polymer(p).get_volume_fraction()}
            for p in range(mol.get_n_polymer_types())]
    return {"model": model, "polymers": pols, "propagators": props, "blocks": blocks}


SLAB_KERNEL = re.compile(r"slab_kernel_\w+<double, 1, 64(, \d+)?>")


def fft_schedule_info(n):
    """The FFT launch shapes the autotuner picked for an n^3 grid in this process
    (slab workgroup threads, pencil kz columns per wavefront, measured candidates)."""
    from langevin_fts_amd import capi

    g = capi.Grid(capi.Context(0), [n] * 3, [1.0] * 3)
    return {p: dict(zip(("slab_threads", "pencil_columns", "note"), g.fft_schedule(code)))
            for p, code in (("fp64", capi.FTS_FP64), ("fp32", capi.FTS_FP32))}


def pmc_child(workload):
    """A short run for the counter passes: one saddle point of two iterations."""
    from langevin_fts_amd import lfts

    w = dict(WORKLOADS[workload], max_iter=2)
    M = int(np.prod(w["nx"]))
    with stdout_to_stderr():
        sim = lfts.LFTS(lfts_params(w, "/tmp/lfts_bench"), random_seed=1)
        rng = np.random.Generator(np.random.PCG64(1))
        sim.saddle_device.set_fields(np.ascontiguousarray(rng.standard_normal((2, M))))
        sim._find_saddle_point_device()


def _counter_means(outdir, counter, name_re):
    vals = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] == counter and name_re.search(row["Kernel_Name"]):
                    vals.append(float(row["Counter_Value"]))
    return sum(vals) / len(vals) if vals else None


def measure_traffic(workload):
    """HBM bytes of one slab launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
    scaled by the same counters on a 256 MiB device copy. None when rocprofv3 is
    missing or a pass fails."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    here = os.path.abspath(__file__)
    env = dict(os.environ, TMPDIR="/tmp")
    out = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            for mode, name_re, extra in (("copy", re.compile(r"copyBuffer"), ["--copy-calibration", "256"]),
                                         ("slab", SLAB_KERNEL, ["--pmc-child", "--workload", workload])):
                d = tempfile.mkdtemp(prefix="fts_pmc_", dir="/tmp")
                cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", mode, "--",
                       sys.executable, here] + extra
                r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                   timeout=180)
                if r.returncode != 0:
                    return None
                out[(counter, mode)] = _counter_means(d, counter, name_re)
                shutil.rmtree(d, ignore_errors=True)
    except (OSError, subprocess.SubprocessError):
        return None
    if any(v is None or v <= 0 for v in out.values()):
        return None
    copy_kib = 256 * 1024
    fetch = out[("FETCH_SIZE", "slab")] * copy_kib / out[("FETCH_SIZE", "copy")]
    write = out[("WRITE_SIZE", "slab")] * copy_kib / out[("WRITE_SIZE", "copy")]
    return {"bytes": (fetch + write) * 1024, "fetch_scale": copy_kib / out[("FETCH_SIZE", "copy")],
            "write_scale": copy_kib / out[("WRITE_SIZE", "copy")]}


def copy_calibration(mib):
    """Plain device copies of a known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import ctypes

    from langevin_fts_amd import capi

    nbytes = mib << 20
    a = capi.DeviceArray(nbytes)
    b = capi.DeviceArray(nbytes)
    a.upload(np.ones(nbytes // 8))
    rt = capi.hip_runtime()
    for _ in range(8):
        if rt.hipMemcpy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), nbytes, 3) != 0:
            raise RuntimeError("hipMemcpy device to device failed")
    rt.hipDeviceSynchronize()
    print(json.dumps({"copy_bytes_per_dispatch": nbytes, "dispatches": 8}), flush=True)


def kernel_roofline(kern, label):
    sl = kern["slab"]
    achieved = sl["bytes"] / (sl["milliseconds"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS, "traffic": None, "kernel": label,
            "timed_launch_count": sl["launches"], "avg_launch_us": 1e3 * sl["milliseconds"] / max(sl["launches"], 1),
            "bytes_per_launch": sl["bytes"] / max(sl["launches"], 1), "timed_launches": f"1 in {SAMPLE}"}


def replica_bench(args, w, lf, lfts):
    """Batching form: w["replicas"] independent runs whose propagators share launches;
    value = replica Langevin steps per second over all replicas."""
    R = w["replicas"]
    M = int(np.prod(w["nx"]))
    params = lfts_params(w, "/tmp/lfts_bench")
    mixed = args.precision == "mixed"
    params["saddle"]["mixed_precision"] = mixed
    with stdout_to_stderr():
        sim = lfts.LFTS(params, random_seed=args.seed, n_replicas=R)
        sim.saddle_device.set_graphs(not args.no_graphs)
        rng = np.random.Generator(np.random.PCG64(args.seed))
        w_aux = np.stack([sim.mpt.to_aux_fields(rng.standard_normal((2, M))) for _ in range(R)])
        dev = sim.saddle_device
        dev.set_fields(np.ascontiguousarray(w_aux))
        dev.set_noise_prev(np.zeros((R, 1, M)))
        dev.find_saddle_point(sim.saddle["tolerance"], sim.saddle["max_iter"], 0)
        for _ in range(args.warmup):
            sim.replica_step()
        # the timed steps run without sampled launch events, so the saddle iterations and
        # Langevin updates replay their captured graphs; a roofline pass of its own follows
        its, fails = [], 0
        dev.field_moments()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            n, _, ok = sim.replica_step()
            its += list(n)
            fails += sum(1 for a in ok if not a)
        dev.field_moments()
        elapsed = time.perf_counter() - t0
        kern = {}
        if not args.no_kernel_timing:
            lf.set_kernel_timing(SAMPLE)
            for _ in range(max(2, args.steps // 4)):
                sim.replica_step()
            kern = {f: lf.get_kernel_timing(f) for f in ("slab", "pencil")}
            lf.set_kernel_timing(0)
    out = {
        "metric": f"L-FTS replica steps/s @{w['nx'][0]}³ ({R} replicas batched)",
        "value": R * args.steps / elapsed, "unit": "replica steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 propagators until err < tol, then f64; f64 Q/phi/AM" if mixed else "f64",
        "data": "synthetic: w_A, w_B ~ N(0,1) from PCG64(seed); replica r's noise from PCG64(seed + r)",
        "config": {"workload": describe(args.workload, w), "replicas": R,
                   "saddle_iterations_per_step": float(np.mean(its)), "saddle_failures": fails,
                   "max_saddle_iterations": int(np.max(its)), "fft_schedule": fft_schedule_info(w["nx"][0]),
                   "iteration_graphs": ("off: launches issued one by one" if args.no_graphs else
                                        "saddle iterations and Langevin updates replayed from captured HIP graphs "
                                        "(the kernel-timed roofline pass runs its launches one by one)")},
    }
    if kern:
        out["roofline"] = kernel_roofline(
            kern, f"slab pass (slab_kernel_wl<double, 1, {w['nx'][0]}, threads>: up to 32 propagators per launch)")
    print(json.dumps(out), flush=True)


def realspace_setup(w):
    """mask (a sphere of radius w["radius"] at the box centre is impenetrable) and
    the grafting field: planes at index 30 of each axis, 1/dx there (devel/RealSpace3D.py:32-78)"""
    nx, lx = w["nx"], w["lx"]
    axes = [np.linspace(-l / 2, l / 2, num=n, endpoint=False) for n, l in zip(nx, lx)]
    xv, yv, zv = np.meshgrid(*axes, indexing="ij")
    mask = np.ones(nx)
    mask[np.sqrt(xv ** 2 + yv ** 2 + zv ** 2) < w["radius"]] = 0.0
    q_init = np.zeros(nx)
    q_init[30, :, :] = 1.0 / (lx[0] / nx[0])
    q_init[:, 30, :] = 1.0 / (lx[1] / nx[1])
    q_init[:, :, 30] = 1.0 / (lx[2] / nx[2])
    return mask.ravel(), q_init.ravel()


def realspace_cpu_baseline(w, budget_s, steps_per_evaluation):
    """The FP64 oracle's real-space steps on the same grid, walls and mask; a
    5-segment chain times the per-step cost, scaled to one evaluation."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    from langevin_fts_amd import langevinfts as lf

    factory = lf.PlatformSelector.create_factory("hip", False)
    mol = factory.create_molecules_information("continuous", w["ds"], {"A": 1.0})
    mol.add_polymer(1.0, [["A", 5 * w["ds"], 0, 1]], {0: "G"})
    an = factory.create_propagator_analyzer(mol, False)
    plan = plan_dict(mol, an, "continuous")
    mask, q_init = realspace_setup(w)
    M = int(np.prod(w["nx"]))
    solver = O.Solver(O.Grid(w["nx"], w["lx"], mask), plan, w["ds"], {"A": 1.0}, realspace_bc=w["bc"])
    n, t0 = 0, time.perf_counter()
    while True:
        solver.compute_propagators({"A": np.zeros(M)}, {"G": q_init})
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t_step = (time.perf_counter() - t0) / (n * 10)
    return {"value": 1.0 / (t_step * steps_per_evaluation), "unit": "evaluations/s", "cores": 1, "kind": "port",
            "sample": f"{10 * n} real-space steps of the same grid on the FP64 C oracle ({1e3 * t_step:.0f} ms each, "
                      f"one thread) x {steps_per_evaluation} steps per evaluation"}


def realspace_bench(args, w, lf):
    """Real-space solver: one step = compute_propagators + compute_concentrations of
    the grafted homopolymer (N = 100, two propagators of 100 Crank-Nicolson steps)
    with w = 0, fields resident in device memory."""
    from langevin_fts_amd import capi

    M = int(np.prod(w["nx"]))
    mask, q_init = realspace_setup(w)
    with stdout_to_stderr():
        factory = lf.PlatformSelector.create_factory("hip", False)
        cb = factory.create_computation_box(w["nx"], w["lx"], bc=w["bc"], mask=mask)
        mol = factory.create_molecules_information("continuous", w["ds"], {"A": 1.0})
        mol.add_polymer(1.0, [["A", 1.0, 0, 1]], {0: "G"})
        an = factory.create_propagator_analyzer(mol, False)
        solver = factory.create_realspace_solver(cb, mol, an)
        d_w = capi.DeviceArray.from_host(np.zeros(M))
        d_q = capi.DeviceArray.from_host(q_init)
        d_phi = capi.DeviceArray(8 * M)

        def step():
            solver.compute_propagators_device({"A": d_w.ptr}, {"G": d_q.ptr})
            solver.compute_concentrations()
            solver.get_total_concentration_device("A", d_phi.ptr)

        for _ in range(args.warmup):
            step()
        solver.get_total_partition(0)   # synchronises the stream
        if not args.no_kernel_timing:
            lf.set_kernel_timing(SAMPLE)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        Q = solver.get_total_partition(0)
        elapsed = time.perf_counter() - t0
        kern = lf.get_kernel_timing("real") if not args.no_kernel_timing else None
        lf.set_kernel_timing(0)
        # the reference's own calling pattern: host fields in, phi out each evaluation
        n_host = max(2, args.steps // 5)
        t0 = time.perf_counter()
        for _ in range(n_host):
            solver.compute_propagators({"A": np.zeros(M)}, q_init={"G": q_init})
            solver.compute_concentrations()
            solver.get_total_concentration("A")
        host_s = (time.perf_counter() - t0) / n_host
    steps_per_eval = sum(e.max_n_segment for e in an.get_computation_propagators().values())
    out = {
        "metric": "real-space propagator evaluations/s @120x100x140 (Crank-Nicolson, walls, nanoparticle mask)",
        "value": args.steps / elapsed, "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: w = 0, grafting planes and a spherical mask as devel/RealSpace3D.py",
        "config": {"workload": f"grafted homopolymer, continuous N=100 (ds=0.01), {w['nx'][0]}x{w['nx'][1]}x"
                               f"{w['nx'][2]} grid, L={w['lx']}, walls {w['bc']}, nanoparticle radius {w['radius']}",
                   "crank_nicolson_steps_per_evaluation": steps_per_eval,
                   "Q": Q, "host_fields_evaluations_per_s": 1.0 / host_s},
    }
    if kern and kern["launches"]:
        achieved = kern["bytes"] / (kern["milliseconds"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                           "kernel": "one Crank-Nicolson step (real_pre_kernel + 3 real_sweep_kernel launches)",
                           "timed_launch_count": kern["launches"],
                           "avg_launch_us": 1e3 * kern["milliseconds"] / kern["launches"],
                           "bytes_per_launch": kern["bytes"] / kern["launches"], "timed_launches": f"1 in {SAMPLE}"}
        out["config"]["step_kernel_share"] = SAMPLE * kern["milliseconds"] * 1e-3 / elapsed
    if args.cpu_seconds > 0:
        with stdout_to_stderr():
            out["cpu_baseline"] = realspace_cpu_baseline(w, args.cpu_seconds, steps_per_eval)
    print(json.dumps(out), flush=True)


def reduce_bench(args, w, lf):
    """Reduce-memory form against the device-resident one on the same evaluation:
    compute_propagators + compute_concentrations + compute_stress of an AB diblock,
    fields N(0, 2^2) resident in device memory. value = reduce-memory evaluations/s."""
    from langevin_fts_amd import capi

    M = int(np.prod(w["nx"]))
    rng = np.random.Generator(np.random.PCG64(args.seed))
    fields = {"A": rng.normal(0.0, 2.0, M), "B": rng.normal(0.0, 2.0, M)}
    res = {}
    with stdout_to_stderr():
        d_w = {m: capi.DeviceArray.from_host(v) for m, v in fields.items()}
        for reduce in (False, True):
            factory = lf.PlatformSelector.create_factory("hip", reduce)
            cb = factory.create_computation_box(w["nx"], w["lx"])
            mol = factory.create_molecules_information(w["model"], w["ds"], {"A": 1.0, "B": 1.0})
            mol.add_polymer(1.0, [["A", w["f"], 0, 1], ["B", 1 - w["f"], 1, 2]])
            an = factory.create_propagator_analyzer(mol, True)
            solver = factory.create_pseudospectral_solver(cb, mol, an)

            def step():
                solver.compute_propagators_device({m: a.ptr for m, a in d_w.items()})
                solver.compute_concentrations()
                solver.compute_stress()   # synchronises the stream

            for _ in range(args.warmup):
                step()
            if reduce and not args.no_kernel_timing:
                lf.set_kernel_timing(SAMPLE)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step()
            elapsed = time.perf_counter() - t0
            kern = None
            if reduce and not args.no_kernel_timing:
                kern = lf.get_kernel_timing("slab")
                lf.set_kernel_timing(0)
            res[reduce] = {"s": elapsed / args.steps, "mem": solver.get_memory_usage(), "kern": kern,
                           "Q": solver.get_total_partition(0)}
            del solver, an, mol, cb, factory
    r, d = res[True], res[False]
    host_bytes = r["mem"][1]
    out = {
        "metric": "propagator evaluations/s @128³, reduce-memory form (propagators in pinned host memory)",
        "value": 1.0 / r["s"], "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * r["s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: w_A, w_B ~ N(0, 2^2) from PCG64(seed)",
        "config": {"workload": f"AB diblock, discrete N=90, f={w['f']}, {w['nx'][0]}^3, L={w['lx'][0]}; one step = "
                               "compute_propagators + compute_concentrations + compute_stress",
                   "device_form_evaluations_per_s": 1.0 / d["s"], "reduce_vs_device": d["s"] / r["s"],
                   "device_bytes": {"device_form": d["mem"][0], "reduce_memory": r["mem"][0]},
                   "pinned_host_bytes": host_bytes,
                   # each propagator slice is written once and read back by the concentration and stress passes
                   "host_link_GBps_lower_bound": 2 * host_bytes / r["s"] / 1e9,
                   "same_Q": r["Q"] == d["Q"]},
    }
    kern = r["kern"]
    if kern and kern["launches"]:
        achieved = kern["bytes"] / (kern["milliseconds"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                           "kernel": "slab pass of the reduce-memory form (q slices stored over the host link)",
                           "timed_launch_count": kern["launches"],
                           "avg_launch_us": 1e3 * kern["milliseconds"] / kern["launches"],
                           "bytes_per_launch": kern["bytes"] / kern["launches"], "timed_launches": f"1 in {SAMPLE}"}
    print(json.dumps(out), flush=True)


def make_solver(lf, platform, w, reduce=False):
    factory = lf.PlatformSelector.create_factory(platform, reduce)
    cb = factory.create_computation_box(w["nx"], w["lx"])
    mol = factory.create_molecules_information(w["model"], w["ds"], w["bonds"])
    for vf, blocks in w["polymers"]:
        mol.add_polymer(vf, blocks)
    an = factory.create_propagator_analyzer(mol, True)
    return factory, cb, mol, an, factory.create_pseudospectral_solver(cb, mol, an)


def baseline_eval_seconds(w, budget_s, with_stress=True):
    """CPU baseline time of one evaluation (propagators, concentrations and, for
    continuous chains, stress) over the real contour: at least one evaluation, more
    while the budget lasts."""
    solver, bonds, cores = baseline_solver(w)
    M = int(np.prod(w["nx"]))
    rng = np.random.default_rng(0)
    fields = {m: w.get("sigma", 1.0) * rng.standard_normal(M) for m in bonds}
    stress = with_stress and w["model"] == "continuous"
    n, t0 = 0, time.perf_counter()
    while True:
        solver.compute_propagators(fields)
        solver.compute_concentrations()
        if stress:
            solver.compute_stress()
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t = (time.perf_counter() - t0) / n
    what = "propagators + concentrations" + (" + stress" if stress else " (no stress)")
    return t, cores, f"{n} evaluation(s) ({what}) of the same grid and polymers, whole contour, {t:.2f} s each; " \
                     + baseline_note(solver, cores)


def eval_bench(args, w, lf):
    """Evaluations/s (compute_propagators + compute_concentrations + compute_stress,
    fields resident in device memory) of one configuration, per precision form."""
    from langevin_fts_amd import capi

    M = int(np.prod(w["nx"]))
    rng = np.random.Generator(np.random.PCG64(args.seed))
    fields = {m: w["sigma"] * rng.standard_normal(M) for m in sorted(w["bonds"])}
    rates, kern, steps_per_eval = {}, None, 0
    with stdout_to_stderr():
        d_w = {m: capi.DeviceArray.from_host(v) for m, v in fields.items()}
        for platform in w.get("precisions", ["hip"]):
            factory, cb, mol, an, solver = make_solver(lf, platform, w)
            steps_per_eval = sum(e.max_n_segment for e in an.get_computation_propagators().values())

            def step():
                solver.compute_propagators_device({m: a.ptr for m, a in d_w.items()})
                solver.compute_concentrations()
                solver.compute_stress()   # synchronises the stream

            for _ in range(args.warmup):
                step()
            first = platform == w.get("precisions", ["hip"])[0]
            if first and not args.no_kernel_timing:
                lf.set_kernel_timing(SAMPLE)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step()
            rates[platform] = args.steps / (time.perf_counter() - t0)
            if first and not args.no_kernel_timing:
                kern = lf.get_kernel_timing("slab")
                lf.set_kernel_timing(0)
            del solver, an, mol, cb, factory
    main_platform = w.get("precisions", ["hip"])[0]
    out = {
        "metric": f"propagator evaluations/s, {w['label']}", "value": rates[main_platform],
        "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 / rates[main_platform], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 propagators/FFTs, f64 Q/phi/stress" if main_platform == "hip-fp32" else "f64",
        "data": f"synthetic: fields N(0, {w['sigma']}^2) from PCG64(seed)",
        "config": {"workload": w["label"] + "; one step = compute_propagators + compute_concentrations + "
                                            "compute_stress", "steps_per_evaluation": steps_per_eval,
                   "evaluations_per_s": rates},
    }
    if len(set(w["nx"])) == 1 and w["nx"][0] in (32, 48, 64, 96, 128):
        out["config"]["fft_schedule"] = fft_schedule_info(w["nx"][0])
    if kern and kern["launches"]:
        achieved = kern["bytes"] / (kern["milliseconds"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBPS, "traffic": None,
                           "kernel": f"slab pass ({main_platform})", "timed_launch_count": kern["launches"],
                           "avg_launch_us": 1e3 * kern["milliseconds"] / kern["launches"],
                           "bytes_per_launch": kern["bytes"] / kern["launches"], "timed_launches": f"1 in {SAMPLE}"}
    if args.cpu_seconds > 0:
        with stdout_to_stderr():
            t, cores, sample = baseline_eval_seconds(w, args.cpu_seconds)
        out["cpu_baseline"] = {"value": 1.0 / t, "unit": "evaluations/s", "cores": cores, "kind": "port",
                               "sample": sample}
    print(json.dumps(out), flush=True)


def c2_fields(w, seed):
    """w_A = -w_B = chiN/2 sum_{j=1..8} cos(k_j . r + phase_j), k_j among the 24 integer
    vectors with |n|^2 = 6 (times 2 pi / L), phases uniform, both from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vecs = sorted({p for v in itertools.permutations((1, 1, 2)) for p in itertools.product(
        *[(a, -a) for a in v])})
    pick = rng.choice(len(vecs), size=8, replace=False)
    phases = rng.uniform(0.0, 2 * np.pi, 8)
    axes = [np.arange(n) * l / n for n, l in zip(w["nx"], w["lx"])]
    x, y, z = np.meshgrid(*axes, indexing="ij")
    field = np.zeros(w["nx"])
    for j, ph in zip(pick, phases):
        k = 2 * np.pi * np.array(vecs[j]) / np.array(w["lx"])
        field += np.cos(k[0] * x + k[1] * y + k[2] * z + ph)
    field *= w["chi_n"] / 2
    return {"A": field, "B": -field}


def scft_bench(args, w, lf):
    """SCFT iterations/s with box relaxation (the driver of examples/scft/scft.py)."""
    from langevin_fts_amd import scft

    params = {"nx": w["nx"], "lx": w["lx"], "box_is_altering": True, "chain_model": w["model"], "ds": w["ds"],
              "segment_lengths": {"A": 1.0, "B": 1.0}, "chi_n": {"A,B": w["chi_n"]},
              "distinct_polymers": [{"volume_fraction": 1.0, "blocks": [{"type": "A", "length": w["f"]},
                                                                        {"type": "B", "length": 1 - w["f"]}]}],
              "optimizer": {"name": "am", "max_hist": 20, "start_error": 1e-2, "mix_min": 0.01, "mix_init": 0.1},
              "max_iter": args.warmup + args.steps, "tolerance": 1e-12}
    fields = c2_fields(w, args.seed)
    with stdout_to_stderr():
        calc = scft.SCFT(params=params)
        # two runs from the same start, W and W + K iterations: their difference is K iterations
        elapsed = []
        for n in (args.warmup, args.warmup + args.steps):
            calc.max_iter = n
            t0 = time.perf_counter()
            calc.run(initial_fields=fields)
            elapsed.append(time.perf_counter() - t0)
    rate = args.steps / (elapsed[1] - elapsed[0])
    out = {
        "metric": f"SCFT iterations/s, {w['label']}", "value": rate, "unit": "iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 / rate, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: w_A = -w_B = chiN/2 sum of 8 cosines with |n|^2 = 6 wave vectors, PCG64(seed) "
                "picks and phases",
        "config": {"workload": w["label"] + "; AM(20, 1e-2, 0.01, 0.1) on [w, L]",
                   "timing": f"run of {args.warmup + args.steps} iterations minus run of {args.warmup} from the same "
                             "start"},
    }
    if args.cpu_seconds > 0:
        wl = dict(w, bonds={"A": 1.0, "B": 1.0}, polymers=[[1.0, [["A", w["f"], 0, 1], ["B", 1 - w["f"], 1, 2]]]],
                  sigma=1.0)
        with stdout_to_stderr():
            tt, cores, sample = baseline_eval_seconds(wl, args.cpu_seconds)
        out["cpu_baseline"] = {"value": 1.0 / tt, "unit": "iterations/s", "cores": cores, "kind": "port",
                               "sample": sample + " (one SCFT iteration = one evaluation; the host-side mixing of "
                                                  "[w, L] is not counted)"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lfts64", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--no-kernel-timing", action="store_true", help="time the steps without sampled launch events")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 8 device-to-device copies of MIB MiB (reference dispatches for HBM counters)")
    ap.add_argument("--pmc-child", action="store_true", help="only run the short pass the counter runs profile")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--guess", default="learned", choices=["learned", "plain"],
                    help="learned: after the plain arm, train the CNN guess and time the steps that use it")
    ap.add_argument("--guess-train-steps", type=int, default=1000, help="Langevin steps the CNN trains on")
    ap.add_argument("--guess-train-iters", type=int, default=48, help="Adam iterations per training step")
    ap.add_argument("--guess-lr", type=float, default=5e-4, help="initial Adam learning rate of the guess")
    ap.add_argument("--guess-lr-final", type=float, default=5e-5, help="learning rate at the end of training")
    ap.add_argument("--guess-replay", type=int, default=64, help="stored training samples replayed in training")
    ap.add_argument("--guess-inputs", default="residual", choices=["residual", "delta", "absolute"],
                    help="residual: the CNN sees w-, the first residual of the search and the starting w+ and "
                         "predicts the correction of w+; delta: w+ += f(w-_new) - f(w-_old)")
    ap.add_argument("--full-run", type=int, default=0, metavar="N",
                    help="lfts64: time one whole N-step run end to end instead of the arms (training included)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="replicas48: issue the saddle iterations' launches one by one instead of replaying graphs")
    ap.add_argument("--guess-use-loss", type=float, default=0.2,
                    help="residual guess: start the training steps' saddle points from it once its recent loss "
                         "is below this (0: only after training)")
    ap.add_argument("--sk-steps", type=int, default=20,
                    help="steps of the untimed structure-function agreement pass (0: skip)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"],
                    help="mixed: saddle iterations on FP32 propagators until the error < tol, then FP64")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.pmc_child:
        pmc_child(args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the L-FTS path here is a "
              "single-GPU build)", file=sys.stderr)
        sys.exit(2)

    # counter passes first, while this process has not initialised the GPU
    traffic = None if args.no_traffic or args.workload != "lfts64" else measure_traffic(args.workload)

    from langevin_fts_amd import langevinfts as lf
    from langevin_fts_amd import lfts

    w = WORKLOADS[args.workload]
    M = int(np.prod(w["nx"]))
    if w.get("replicas", 1) > 1:
        replica_bench(args, w, lf, lfts)
        return
    if w.get("realspace"):
        realspace_bench(args, w, lf)
        return
    if w.get("kind") == "eval":
        eval_bench(args, w, lf)
        return
    if w.get("kind") == "scft":
        scft_bench(args, w, lf)
        return
    if w.get("reduce"):
        reduce_bench(args, w, lf)
        return
    params = lfts_params(w, "/tmp/lfts_bench")
    mixed = args.precision == "mixed"
    params["saddle"]["mixed_precision"] = mixed
    params["learned_guess"] = {"train_steps": args.guess_train_steps, "lr": args.guess_lr,
                               "lr_final": args.guess_lr_final, "train_iters": args.guess_train_iters,
                               "replay": args.guess_replay, "inputs": args.guess_inputs,
                               "use_loss": args.guess_use_loss, "seed": 0}
    with stdout_to_stderr():
        sim = lfts.LFTS(params, random_seed=args.seed)
        tol = sim.saddle["tolerance"]

        def precision(low):
            if mixed:
                sim.saddle_device.set_low_precision_solver(sim.low_precision_solver if low else None, tol)

        precision(False)
        rng = np.random.Generator(np.random.PCG64(args.seed))
        w_mono = np.stack([rng.standard_normal(M), rng.standard_normal(M)])
        sim.saddle_device.set_fields(np.ascontiguousarray(sim.mpt.to_aux_fields(w_mono)))
        sim.saddle_device.set_noise_prev(np.zeros([len(sim.mpt.aux_fields_real_idx), M]))
        sim._find_saddle_point_device()
        warm = {"iterations": 0, "failures": 0}
        langevin_steps(sim, args.warmup, warm)

        full = full_run(args, sim, precision, w) if args.full_run else None
    if full is not None:
        print(json.dumps(full), flush=True)
        return
    with stdout_to_stderr():
        if args.guess == "learned":
            # online training on the run's own saddle points (plain saddle points, then Adam steps)
            precision(True)
            train = {"iterations": 0, "failures": 0}
            t0 = time.perf_counter()
            langevin_steps(sim, args.guess_train_steps, train, learned=True)
            sim.saddle_device.field_moments()
            train_s = time.perf_counter() - t0
            precision(False)

        # plain arm: the reference's method (each saddle point starts from the previous w+);
        # every arm runs without sampled launch events, which go to a separate pass below
        plain = {"iterations": 0, "failures": 0}
        sim.saddle_device.field_moments()   # synchronises the stream
        t0 = time.perf_counter()
        langevin_steps(sim, args.steps, plain)
        sim.saddle_device.field_moments()
        plain_s = time.perf_counter() - t0

        def arm(learned_guess, low):
            precision(low)
            st = {"iterations": 0, "failures": 0}
            t0 = time.perf_counter()
            langevin_steps(sim, args.steps, st, learned=learned_guess)
            sim.saddle_device.field_moments()
            st["s"] = time.perf_counter() - t0
            precision(False)
            return st

        # the same tolerance, each arm continuing from the state the previous one left
        learned = learned64 = plain_mixed = None
        if args.guess == "learned":
            learned64 = arm(True, False)    # saddle points from the CNN's prediction, FP64 iterations
        if mixed:
            plain_mixed = arm(False, True)  # the reference's start, FP32 iterations until err < tol
        if args.guess == "learned" and mixed:
            learned = arm(True, True)       # both: the CNN's start and FP32 iterations until err < tol
            learned_s = learned["s"]

        # roofline pass: plain FP64 steps with one launch in SAMPLE timed by HIP events
        kern = {}
        roof_s = None
        if not args.no_kernel_timing:
            lf.set_kernel_timing(SAMPLE)
            roof = {"iterations": 0, "failures": 0}
            sim.saddle_device.field_moments()
            t0 = time.perf_counter()
            langevin_steps(sim, args.steps, roof)
            sim.saddle_device.field_moments()
            roof_s = time.perf_counter() - t0
            kern = {f: lf.get_kernel_timing(f) for f in ("slab", "pencil")}
            lf.set_kernel_timing(0)

        # structure-function agreement (untimed): from one state and with one noise
        # sequence, the plain and learned-guess FP64 arms each run n_sf steps
        sk = None
        if args.guess == "learned" and args.sk_steps > 0:
            sk = structure_function_agreement(sim, args.sk_steps, args.seed)

    # value: the learned guess with FP64 saddle iterations (the reference's precision)
    stats = learned64 if learned64 is not None else plain
    elapsed = learned64["s"] if learned64 is not None else plain_s
    steps_per_s = args.steps / elapsed
    iters = stats["iterations"] / args.steps
    plain_iters = plain["iterations"] / args.steps
    out = {
        "metric": "L-FTS steps/s @64³ (learned vs plain AM guess, same tol); propagator HBM GB/s",
        "value": steps_per_s, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: w_A, w_B ~ N(0,1) from PCG64(seed); Langevin noise from numpy PCG64",
        "config": {"workload": describe(args.workload, w),
                   "guess": (f"learned CNN ({args.guess_inputs} form) + Anderson mixing, FP64 saddle iterations"
                             if args.guess == "learned" else "plain Anderson mixing, FP64"),
                   "saddle_iterations_per_step": iters, "saddle_failures": stats["failures"],
                   "max_saddle_error": stats.get("max_error"),
                   "plain_steps_per_s": args.steps / plain_s, "plain_saddle_iterations_per_step": plain_iters,
                   "plain_saddle_failures": plain["failures"],
                   "propagator_evaluations_per_s": plain["iterations"] / plain_s},
    }
    out["config"]["fft_schedule"] = fft_schedule_info(w["nx"][0])
    if learned64 is not None:
        out["config"]["learned_fp64_steps_per_s"] = args.steps / learned64["s"]
        out["config"]["learned_fp64_saddle_iterations_per_step"] = learned64["iterations"] / args.steps
    if plain_mixed is not None:
        out["config"]["plain_mixed_steps_per_s"] = args.steps / plain_mixed["s"]
        out["config"]["plain_mixed_saddle_iterations_per_step"] = plain_mixed["iterations"] / args.steps
    if learned is not None and mixed:
        # FP32 propagators while the error >= tol; the converging iteration (and the
        # last permitted one) is always FP64, so max_saddle_error is an FP64 error
        out["config"]["mixed_precision_arm"] = {
            "steps_per_s": args.steps / learned_s, "saddle_iterations_per_step": learned["iterations"] / args.steps,
            "saddle_failures": learned["failures"], "max_saddle_error": learned.get("max_error"),
            "vs_plain_fp64": (args.steps / learned_s) / (args.steps / plain_s),
            "dtype": "f32 propagators until err < tol, then f64; f64 Q/phi/AM"}
    if sk is not None:
        out["config"]["structure_function_agreement"] = sk
    if args.guess == "learned":
        n_train, n_run = args.guess_train_steps, 10000
        train_step_s = train_s / max(n_train, 1)
        out["config"].update({
            "learned_vs_plain": (args.steps / learned64["s"]) / (args.steps / plain_s),
            "guess_train_steps": n_train, "train_step_ms": 1e3 * train_step_s,
            "train_saddle_iterations_per_step": train["iterations"] / max(n_train, 1),
            # the newest sample's loss (scaled units) over the first and the last 50 training steps
            "train_loss_first50": float(np.mean(sim.guess_losses[:50])) if sim.guess_losses else None,
            "train_loss_last50": float(np.mean(sim.guess_losses[-50:])) if sim.guess_losses else None,
            # a 10,000-step run that trains on its first n_train steps, at the measured rates
            "run10k_steps_per_s": n_run / (n_train * train_step_s + (n_run - n_train) * learned64["s"] / args.steps)})
    if kern:
        sl, pe = kern["slab"], kern["pencil"]
        achieved = sl["bytes"] / (sl["milliseconds"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBPS,
                           "traffic": traffic["bytes"] if traffic else None,
                           "kernel": "slab pass (slab_kernel_wl<double, 1, 64, threads>: 2 propagators per "
                                     "launch), timed in a pass of plain FP64 steps after the arms",
                           "timed_launch_count": sl["launches"],
                           "avg_launch_us": 1e3 * sl["milliseconds"] / max(sl["launches"], 1),
                           "bytes_per_launch": sl["bytes"] / max(sl["launches"], 1),
                           "timed_launches": f"1 in {SAMPLE}"}
        if traffic:
            out["roofline"]["traffic_counter_scale"] = {"FETCH_SIZE": traffic["fetch_scale"],
                                                        "WRITE_SIZE": traffic["write_scale"]}
        both_ms = sl["milliseconds"] + pe["milliseconds"]
        out["config"]["propagator_GBps"] = (sl["bytes"] + pe["bytes"]) / (both_ms * 1e-3) / 1e9
        out["config"]["pencil_GBps"] = pe["bytes"] / (pe["milliseconds"] * 1e-3) / 1e9
        out["config"]["pencil_avg_launch_us"] = 1e3 * pe["milliseconds"] / max(pe["launches"], 1)
        # kernel time of the roofline pass; the timed launches are one in SAMPLE of each family
        out["config"]["propagator_kernel_share"] = SAMPLE * both_ms * 1e-3 / roof_s
    if args.cpu_seconds > 0:
        with stdout_to_stderr():
            out["cpu_baseline"] = cpu_baseline(w, plain_iters, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
