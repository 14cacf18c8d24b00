"""Build the in-tree native libraries of langevin_fts_amd.

    libfts_hip.so          HIP kernels + the C ABI of include/fts_hip.h (hipcc, gfx950)
    langevinfts*.so        pybind11 module: the reference's Python surface over the C ABI

Run with `python -m langevin_fts_amd.build` (or through __graft_entry__.build()).
Objects are cached in langevin_fts_amd/_build/ and rebuilt when a source or
header is newer than the object.
"""
import glob
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
HIP_DIR = os.path.join(PKG, "csrc", "hip")
HOST_DIR = os.path.join(PKG, "csrc", "host")
BUILD = os.path.join(PKG, "_build")
ARCH = os.environ.get("FTS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    # keep the compiler from emitting packed FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
] + os.environ.get("FTS_HIP_DEFINES", "").split()   # tuning defines, e.g. -DFTS_STREAM_STORES=1


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build_hip(force=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    headers = glob.glob(os.path.join(HIP_DIR, "*.hpp")) + [os.path.join(REPO, "include", "fts_hip.h")]
    sources = sorted(glob.glob(os.path.join(HIP_DIR, "*.hip")))
    objs = []
    todo = []
    for src in sources:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        if force or _newer(obj, [src] + headers):
            todo.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, src])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    lib = os.path.join(PKG, "libfts_hip.so")
    if force or todo or _newer(lib, objs):
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib] + objs)
    return lib


def build_module(force=False):
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "langevinfts" + suffix)
    sources = sorted(glob.glob(os.path.join(HOST_DIR, "*.cpp")))
    headers = glob.glob(os.path.join(HOST_DIR, "*.h")) + [os.path.join(REPO, "include", "fts_hip.h")]
    if not sources:
        return None
    lib = os.path.join(PKG, "libfts_hip.so")
    if not (force or _newer(out, sources + headers + [lib])):
        return out
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(REPO, "include"), HOST_DIR]
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", "-Wall"]
    for i in inc:
        cmd += ["-I", i]
    cmd += ["-o", out] + sources + ["-L", PKG, "-lfts_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    return out


def build(force=False):
    lib = build_hip(force=force)
    mod = build_module(force=force)
    return lib, mod


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
