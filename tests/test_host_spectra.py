"""Host-side bookkeeping of the mixed path's V list (CPU only).

* The packed half spectrum, its per-line form, the packing of translation
  tables and the table checks (symmetric 3 x 3 DOF block, mirror symmetry of the
  wrap-free tables): tests/host_spectra_check.cpp, built with hipcc against the
  in-tree library and run here.
* Why the wrap-free table changes no result: for the FFT block convolution of
  the V list (sources and targets on the grid indices [np/2, 3np/2) of the
  periodic n = 2 np grid, reference layout fmm3d_mpi/matmgnt3d_mpi.cpp:276-431),
  the kernel entries of grid index np (displacement -np) never reach a target,
  so zeroing them leaves the target block unchanged; and with them zeroed the
  spectra of opposite offsets are conjugates for an even kernel, which they are
  not with them.  Checked here on numpy FFTs, independently of the library.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fmm-g_amd", "lib", "libkifmm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_spectrum_bookkeeping(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libkifmm.so is not built")
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc is not available")
    exe = str(tmp_path / "host_spectra_check")
    src = os.path.join(ROOT, "tests", "host_spectra_check.cpp")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
           "-fopenmp", src, "-o", exe, "-L" + os.path.dirname(LIB), "-lkifmm", "-Wl,-rpath," + os.path.dirname(LIB)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-3000:] + r.stderr[-1000:]


def _kernel_grid(np_, s, wrap):
    """K(g - 2 s) = 1 / |g - 2 s| on the regular grid (index i -> displacement
    i for i < np, i - n above), array [x][y][z]; wrap=False zeroes index np."""
    n = 2 * np_
    step = 2.0 / (np_ - 1)
    i = np.arange(n)
    g = np.where(i < np_, i, i - n) * step
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt((X - 2.0 * s[0]) ** 2 + (Y - 2.0 * s[1]) ** 2 + (Z - 2.0 * s[2]) ** 2)
    K = 1.0 / r
    if not wrap:
        K[np_, :, :] = 0.0
        K[:, np_, :] = 0.0
        K[:, :, np_] = 0.0
    return K


@pytest.mark.parametrize("np_", [4, 6])
def test_wrap_entries_reach_no_target(np_):
    n, o = 2 * np_, np_ // 2
    rng = np.random.default_rng(np_)
    x = np.zeros((n, n, n))
    x[o:o + np_, o:o + np_, o:o + np_] = rng.standard_normal((np_, np_, np_))
    X = np.fft.fftn(x)
    for s in [(2, 0, 0), (3, -1, 2), (-2, -3, 1)]:
        full = np.real(np.fft.ifftn(np.fft.fftn(_kernel_grid(np_, s, True)) * X))
        free = np.real(np.fft.ifftn(np.fft.fftn(_kernel_grid(np_, s, False)) * X))
        blk = (slice(o, o + np_),) * 3
        ref = np.abs(full[blk]).max()
        assert np.abs(full[blk] - free[blk]).max() <= 1e-13 * ref
        # the entries do matter outside the target block
        assert np.abs(full - free).max() > 1e-3 * ref


@pytest.mark.parametrize("np_", [4, 6])
def test_wrap_free_spectra_of_opposite_offsets_are_conjugates(np_):
    for s in [(2, 0, 0), (3, -1, 2), (2, 2, -3)]:
        m = tuple(-v for v in s)
        a, b = np.fft.fftn(_kernel_grid(np_, s, False)), np.fft.fftn(_kernel_grid(np_, m, False))
        assert np.abs(a - np.conj(b)).max() <= 1e-12 * np.abs(a).max()
        a, b = np.fft.fftn(_kernel_grid(np_, s, True)), np.fft.fftn(_kernel_grid(np_, m, True))
        assert np.abs(a - np.conj(b)).max() > 1e-3 * np.abs(a).max()
