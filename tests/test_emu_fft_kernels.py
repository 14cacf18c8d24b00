"""The generic FFT pass kernels of fts_fft.hpp (slab_kernel, xpass_kernel), built
for the host from their own source by tests/emu/fft_block_emu.cpp and run as
emulated workgroups (one std::thread per GPU thread, barriers that report a
deadlock), with AddressSanitizer on the LDS array and the global buffers and
never-written LDS filled with NaN. Each grid's forward transform (slab FWD, then the
x pass) and inverse (x pass, then slab INV1) must equal numpy's rfftn / irfftn.

The 7x11x13 FP32 case is the transform that hung on the MI355X (the library refuses
FP32 on grids with prime factors above 5, and its GPU tests are skipped): here every
LDS index stays in range, every thread reaches the same barriers and the result is
right, for the same launch geometry the library uses. CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(REPO, "tests", "emu")
HIP = os.path.join(REPO, "langevin_fts_amd", "csrc", "hip")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(EMU),
                                reason="needs g++ and tests/emu")


def choose_threads(work):
    return max(64, min(1024, -(-(-(-work // 8)) // 64) * 64))


def lds_bytes(kind, n, esize):
    """the kernels' dynamic LDS (Grid::slab_lds / Grid::x_lds) for the grid"""
    nx, ny, nz = n
    nzc = nz // 2 + 1
    if kind == "slab":
        rows = 2 * ((ny + 1) // 2)
        return (rows * nzc + ny + nz) * 2 * esize
    C, tiles = nzc, 1
    while ny * tiles < 128 and C > 8:
        C = -(-C // 2)
        tiles = -(-nzc // C)
    while (nx * C + nx) * 16 > 64 * 1024 and C > 1:
        C = -(-C // 2)
        tiles = -(-nzc // C)
    return (nx * C + nx) * 2 * esize


def build(tmp, kind, n, esize):
    exe = os.path.join(tmp, f"emu_{kind}_{'x'.join(map(str, n))}_{esize}")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
           f"-DLDS_BYTES={lds_bytes(kind, n, esize)}", "-I", EMU, "-I", HIP, "-o", exe,
           os.path.join(EMU, "fft_block_emu.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def run(exe, kind, direction, esize, n, a, tmp):
    fi, fo = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    a.tofile(fi)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe, kind, direction, str(esize)] + [str(x) for x in n] + [fi, fo], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(fo, dtype=a.dtype), r.stdout


@pytest.mark.parametrize("n,esize", [((7, 11, 13), 4), ((7, 11, 13), 8), ((13, 6, 10), 4), ((5, 4, 3), 4)])
def test_emulated_kernels_transform_correctly(tmp_path, n, esize):
    tmp = str(tmp_path)
    dt = np.float32 if esize == 4 else np.float64
    rng = np.random.default_rng(3)
    a = rng.standard_normal(n)
    slab, xpass = build(tmp, "slab", n, esize), build(tmp, "xpass", n, esize)
    s1, log1 = run(slab, "slab", "fwd", esize, n, a.astype(dt).ravel(), tmp)
    s2, log2 = run(xpass, "xpass", "fwd", esize, n, s1, tmp)
    spec = s2.astype(np.float64).view(np.complex128) if esize == 8 else \
        (s2[0::2].astype(np.float64) + 1j * s2[1::2].astype(np.float64))
    ref = np.fft.rfftn(a).ravel()
    tol = 1e-12 if esize == 8 else 2e-6
    err = np.max(np.abs(spec - ref)) / np.max(np.abs(ref))
    print(n, esize, log1.splitlines()[0], "|", log2.splitlines()[0], f"forward rel err {err:.1e}")
    assert err < tol
    b1, _ = run(xpass, "xpass", "inv", esize, n, s2, tmp)
    b2, _ = run(slab, "slab", "inv", esize, n, b1, tmp)
    err = np.max(np.abs(b2.astype(np.float64) - a.ravel())) / np.max(np.abs(a))
    assert err < tol
