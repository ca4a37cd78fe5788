"""The tt0-style driver (fmm-g_amd/bin/kifmm_tt0, the reference's tt0.cpp on one
GPU over the C ABI).

CPU: with -kifmm_host_only it builds the tree without a GPU; its source
coordinates (-dump_src_coord) must be the reference driver's on rank 0: drand48
seeded with 0 (a 48-bit linear congruential generator, restated here from its
published constants), the distribution's formula, then Morton order at depth 30
with duplicates removed and each point moved to its cell centre
(tt0.cpp:153, 203-326).  GPU: the full run prints "Relative error is: %e"
(tt0.cpp:557), checked against an FMM-accuracy bound.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TT0 = os.path.join(ROOT, "fmm-g_amd", "bin", "kifmm_tt0")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Drand48:
    """POSIX drand48: X <- (0x5DEECE66D X + 0xB) mod 2^48, value X / 2^48;
    srand48(s) sets X = s * 2^16 + 0x330E."""

    def __init__(self, seed):
        self.x = ((seed & 0xFFFFFFFF) << 16) | 0x330E

    def __call__(self):
        self.x = (0x5DEECE66D * self.x + 0xB) & ((1 << 48) - 1)
        return self.x / float(1 << 48)


def morton_dedup(p):
    L = 30
    c = np.floor(np.ldexp(p, L)).astype(np.int64)

    def key(x, y, z):
        k = 0
        for b in range(L - 1, -1, -1):
            k = (k << 3) | (((z >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((x >> b) & 1)
        return k
    keys = sorted({key(int(x), int(y), int(z)): (x, y, z) for x, y, z in c}.items())
    cells = np.array([v for _, v in keys], dtype=np.float64)
    return np.ldexp(cells + 0.5, -L)


def run(args, timeout=600):
    return subprocess.run([TT0] + args, capture_output=True, text=True, timeout=timeout)


def read_dump(path):
    with open(path) as fh:
        vals = [float(t) for t in fh.read().split()[3:-1]]
    return np.array(vals).reshape(-1, 3)


@pytest.mark.parametrize("dist", ["uniform", "tt1", "normal", "logNormal", "sphereUniformAngles"])
def test_driver_points_are_the_reference_drivers(tmp_path, dist):
    n = 3000
    dump = str(tmp_path / "src.m")
    r = run(["-distribution", dist, "-numsrc", str(n), "-preload_numsrc", "10", "-kt", "111",
             "-fmm3d_np", "4", "-fmm3d_let3d_ptsMax", "32", "-fmm3d_let3d_maxLevel", "30",
             "-trgs_and_srcs_coinc", "-kifmm_host_only", "-dump_src_coord", dump])
    assert r.returncode == 0, r.stderr
    assert "*** Dry run (preloading) ***" in r.stdout and "*** Final run ***" in r.stdout
    got = read_dump(dump)
    rnd = Drand48(0)
    R = 0.25
    lc = [0.5 + R * (2.0 * rnd() - 1.0) for _ in range(3)]
    if dist == "uniform":
        p = np.array([rnd() for _ in range(3 * n)]).reshape(n, 3)
    elif dist == "tt1":
        p = np.array([[lc[d] + (0.499 - R) * (2.0 * rnd() - 1.0) for d in range(3)] for _ in range(n)])
    elif dist == "sphereUniformAngles":
        rows = []
        for _ in range(n):
            phi, th = 2 * np.pi * rnd(), np.pi * rnd()
            rows.append([0.5 + 0.49 * np.sin(th) * np.cos(phi), 0.5 + 0.49 * np.sin(th) * np.sin(phi),
                         0.5 + 0.49 * np.cos(th)])
        p = np.array(rows)
    else:
        cache = []

        def gaussian(mean, sd):
            if cache:
                return mean + sd * cache.pop(0)
            while True:
                x1, x2, x3 = (2 * rnd() - 1 for _ in range(3))
                rr = x1 * x1 + x2 * x2 + x3 * x3
                if rr < 1:
                    break
            rr = np.sqrt(-2.0 * np.log(rr) / rr)
            cache.extend([rr * x2, rr * x3])
            return mean + sd * rr * x1
        rows = []
        while len(rows) < n:
            if dist == "normal":
                t = [gaussian(0.5, 0.16) for _ in range(3)]
                ok = all(0 <= v < 1.0 for v in t)
            else:
                t = [np.exp(gaussian(-1.75, 1.0)) for _ in range(3)]
                ok = all(v < 1.0 for v in t)
            if ok:
                rows.append(t)
        p = np.array(rows)
    want = morton_dedup(p)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_driver_errors():
    r = run(["-numsrc", "10"])
    assert r.returncode != 0 and "-distribution" in r.stderr
    r = run(["-distribution", "cube", "-numsrc", "10", "-preload_numsrc", "10", "-kt", "111", "-fmm3d_np", "4",
             "-fmm3d_let3d_ptsMax", "32", "-fmm3d_let3d_maxLevel", "30", "-trgs_and_srcs_coinc", "-kifmm_host_only"])
    assert r.returncode != 0 and "Invalid distribution" in r.stderr


@pytest.mark.gpu
def test_driver_full_run_reports_error():
    """The options file of a reference-style run: 20,000 uniform sources =
    targets, Laplace np 4, -do_2nd_eval, 600 checked targets (all of them in
    the dry run, -preload_check_all).  np 4 reaches about 1e-4."""
    r = run(["-options_file", os.path.join(GOLDEN, "options_tt0_uniform")])
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    errs = [float(m) for m in re.findall(r"Relative error is: (\S+)", r.stdout)]
    assert len(errs) == 2
    assert all(0 < e < 1e-3 for e in errs), errs
    assert "FMM evaluate (2nd)" in r.stdout and "EvlUList" in r.stdout
