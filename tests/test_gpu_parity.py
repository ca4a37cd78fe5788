"""GPU parity: the HIP evaluate pass through the C ABI against the CPU oracle.

Bar (FP64 path): relative 2-norm difference over all targets <= 1e-12 against the
oracle's restatement of the reference CPU evaluate (fmm3d_eval_mpi.cpp:396-1515)
run on the same tree with the same pseudo-inverse tables; plus the reference's
own accuracy check against the direct sum (fmm3d_check_mpi.cpp:159-165).
"""
import numpy as np
import pytest

import kifmm
import oracle as O
from kifmm import workloads as W

pytestmark = pytest.mark.gpu

PARITY = 1e-12


def rel(a, b):
    r = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("relative difference %.3e" % r, flush=True)
    return r


def run_pair(kt, coefs, np_, pts_max, pos, nor, den, max_level=30, tpos=None):
    print("case kt=%d np=%d n=%d ptsMax=%d" % (kt, np_, len(pos), pts_max), flush=True)
    f = kifmm.FMM3d(kt, coefs, np_, pts_max, max_level)
    f.set_points(pos, nor, tpos)
    f.setup()
    val = f.evaluate(den)
    km = O.mm_kernel(kt)
    sd, td = O.src_dof(km), O.trg_dof(km)
    nue = np_ ** 3 - (np_ - 2) ** 3
    nuc = (np_ + 2) ** 3 - np_ ** 3
    pu = f.operator(0).reshape(nuc * td, nue * sd).T
    pd = f.operator(1).reshape(nue * td, nue * sd).T
    mmc = list(coefs) if kt not in (312, 322) else [1.0, 0.25]
    ops = O.Ops(km, mmc, np_, pu, pd)
    tree = O.Tree(pos, tpos, pts_max=pts_max, max_level=max_level)
    ref = O.evaluate(tree, ops, kt, coefs, pos, nor, pos if tpos is None else tpos, den)
    return f, val, ref


def test_laplace_uniform_32k():
    """32,768 uniform points, Laplace single layer, np 4, ptsMax 64 (the CPU-runnable case)."""
    n = 32768
    pos = W.uniform(n, W.seed(0, 0))
    den = W.densities(n, 1, W.seed(0, 1))
    f, val, ref = run_pair(111, [1.0, 0.25], 4, 64, pos, np.zeros_like(pos), den)
    assert rel(val, ref) <= PARITY
    err = f.check(den, val, 512, seed=W.seed(0, 3))
    assert err < 5e-4, err


@pytest.mark.parametrize("kt,coefs", [(121, [1.0, 0.25]), (311, [1.0, 0.25]), (312, [1.0, 0.25]),
                                      (321, [1.0, 0.25]), (322, [1.0, 0.25]), (511, [1.0, 0.3]),
                                      (521, [1.0, 0.3])])
def test_kernels_spheres(kt, coefs):
    """Every supported user kernel on adaptive sphere-surface points (W and X lists active)."""
    n = 6000
    pos, nor = W.spheres(n, 7 + kt, nsph=6)
    sd = O.src_dof(kt)
    den = W.densities(n, sd, 11 + kt, kind="normal")
    f, val, ref = run_pair(kt, coefs, 4, 24, pos, nor, den)
    assert rel(val, ref) <= PARITY


def test_laplace_np6_gaussians():
    """Adaptive Gaussian clusters with a depth cap, np 6 (all list types and the depth cap)."""
    n = 40000
    pos = W.gaussians(n, 99)
    den = W.densities(n, 1, 100)
    f, val, ref = run_pair(111, [1.0, 0.25], 6, 32, pos, np.zeros_like(pos), den, max_level=7)
    assert rel(val, ref) <= PARITY
    err = f.check(den, val, 1024, seed=5)
    assert err < 1e-5, err


def test_navier_np8():
    n = 5000
    pos = W.uniform(n, 31)
    den = W.densities(n, 3, 32, kind="normal")
    f, val, ref = run_pair(511, [1.0, 0.3], 8, 60, pos, np.zeros_like(pos), den, max_level=10)
    assert rel(val, ref) <= PARITY


def test_separate_targets():
    """Targets distinct from sources (trgData on its own point set)."""
    pos = W.uniform(8000, 41)
    tpos = W.uniform(3000, 42)
    den = W.densities(8000, 1, 43)
    f, val, ref = run_pair(111, [1.0, 0.25], 4, 32, pos, np.zeros_like(pos), den, tpos=tpos)
    assert rel(val, ref) <= PARITY


def test_known_answer_single_level():
    """N <= ptsMax: the root split is forced, every leaf's U list holds all 8
    siblings, so the FMM equals the direct sum to round-off."""
    n = 50
    pos = W.uniform(n, 51)
    den = W.densities(n, 1, 52)
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 64, 30)
    f.set_points(pos)
    f.setup()
    val = f.evaluate(den)
    d = O.direct(111, [1.0, 0.25], pos, np.zeros_like(pos), den, pos)
    assert rel(val, d) < 1e-14


def test_empty_leaves_and_duplicates():
    """Coincident points (r < 1e-8 excluded) and empty octants."""
    base = W.uniform(300, 61)
    pos = np.concatenate([base, base[:40], np.full((5, 3), 0.123)])
    den = W.densities(len(pos), 1, 62)
    f, val, ref = run_pair(111, [1.0, 0.25], 4, 8, pos, np.zeros_like(pos), den)
    assert rel(val, ref) <= PARITY


def test_repeat_evaluate_and_linearity():
    """evaluate() is repeatable (buffers re-zeroed) and linear in the densities.

    Linearity holds to round-off times the amplification of the np=6
    pseudo-inverses (cutoff 1e-7): the CPU oracle differs from itself by
    1e-11..4e-11 when only its FMA contraction changes, so the bound is 1e-10."""
    n = 20000
    pos = W.uniform(n, 71)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30)
    f.set_points(pos)
    f.setup()
    a, b = W.densities(n, 1, 72), W.densities(n, 1, 73)
    va, vb = f.evaluate(a), f.evaluate(b)
    vab = f.evaluate(2.0 * a - 0.5 * b)
    assert np.array_equal(f.evaluate(a), va)
    assert rel(vab, 2.0 * va - 0.5 * vb) < 1e-10


def test_device_pointers(gpu):
    import torch
    n = 16384
    pos = W.uniform(n, 81)
    den = W.densities(n, 1, 82)
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 64, 30)
    f.set_points(pos)
    f.setup()
    host = f.evaluate(den)
    d = torch.from_numpy(den).to(gpu)
    v = torch.empty(n, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    f.evaluate_device(d, v)
    torch.cuda.synchronize()
    assert np.array_equal(v.cpu().numpy(), host)


@pytest.mark.parametrize("kt,np_,coefs,gen,tol", [(111, 6, [1.0, 0.25], "gauss", 3e-7),
                                                  (111, 4, [1.0, 0.25], "uniform", 3e-7),
                                                  (321, 4, [1.0, 0.25], "spheres", 3e-5),
                                                  (511, 6, [1.0, 0.3], "uniform", 1e-6)])
def test_mixed_precision(kt, np_, coefs, gen, tol):
    """Mixed path (FP32 near field, S2M/S2L and V-list spectra; FP64 pseudo-inverse
    applies and final sums): close to the FP64 path and no less accurate
    against the direct sum beyond that difference."""
    n = 20000
    if gen == "spheres":
        pos, nor = W.spheres(n, 91, nsph=6)
    else:
        pos = W.gaussians(n, 92) if gen == "gauss" else W.uniform(n, 93)
        nor = np.zeros_like(pos)
    den = W.densities(n, O.src_dof(kt), 94, kind="normal")
    print("case kt=%d np=%d mixed" % (kt, np_), flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision=prec)
        f.set_points(pos, nor)
        f.setup()
        out[prec] = f.evaluate(den)
        out[prec + "_err"] = f.check(den, out[prec], 2000, seed=7)
    d = rel(out["mixed"], out["fp64"])
    print("errors vs direct: fp64 %.3e mixed %.3e" % (out["fp64_err"], out["mixed_err"]), flush=True)
    assert d < tol
    assert out["mixed_err"] <= out["fp64_err"] + 2 * d


@pytest.mark.parametrize("kt,np_,coefs,tol", [(111, 3, [1.0, 0.25], 1e-6), (111, 5, [1.0, 0.25], 1e-6),
                                              (111, 7, [1.0, 0.25], 1e-6), (111, 8, [1.0, 0.25], 1e-6),
                                              (511, 8, [1.0, 0.3], 2e-6)])
def test_mixed_every_np(kt, np_, coefs, tol):
    """The mixed path's packed half spectrum at the expansion orders the other
    tests leave out: the transforms handle the interior-plane and the
    mirrored-plane lines in lane rounds of their own (both rounds loaded
    together up to np 6, one round at a time above), and the Hadamard sweeps
    rows padded to a multiple of 16 frequencies.  Mixed against FP64 on the same
    tree, and no less accurate against the direct sum beyond that difference."""
    n = 8000
    pos = W.gaussians(n, 95)
    den = W.densities(n, O.src_dof(kt), 96, kind="normal")
    print("case kt=%d np=%d mixed vs fp64" % (kt, np_), flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(kt, coefs, np_, 32, 30, precision=prec)
        f.set_points(pos)
        f.setup()
        out[prec] = f.evaluate(den)
        out[prec + "_err"] = f.check(den, out[prec], 1000, seed=11)
    d = rel(out["mixed"], out["fp64"])
    print("  mixed vs fp64 %.3e, errors vs direct: fp64 %.3e mixed %.3e" % (d, out["fp64_err"], out["mixed_err"]),
          flush=True)
    assert d < tol
    assert out["mixed_err"] <= out["fp64_err"] + 2 * d


def test_mixed_near_cut_across_leaves():
    """The mixed near field drops the 1e-8 pair test where the host proves no
    pair can need it (only the target leaf's own sources can).  Points 5e-9
    apart on either side of the plane x = 1/2 (a leaf face at every level)
    must still have their pair excluded, as the FP64 path (test on every pair)
    does; a missed cut would add terms of order 1e7."""
    rng = np.random.default_rng(95)
    base = W.uniform(6000, 96)
    yz = rng.random((6, 2)) * 0.8 + 0.1
    left = np.column_stack([np.full(6, 0.5 - 2.5e-9), yz])
    right = np.column_stack([np.full(6, 0.5 + 2.5e-9), yz])
    pos = np.concatenate([base, left, right, base[:20]])  # plus duplicates in one leaf
    den = W.densities(len(pos), 1, 97)
    print("case mixed near cut n=%d" % len(pos), flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30, precision=prec)
        f.set_points(pos)
        f.setup()
        out[prec] = f.evaluate(den)
    assert np.all(np.isfinite(out["mixed"]))
    assert rel(out["mixed"], out["fp64"]) < 3e-7
    k = slice(6000, 6012)
    assert rel(out["mixed"][k], out["fp64"][k]) < 1e-5


def test_mixed_leaf_split():
    """The mixed near field cuts each leaf's targets into chunks sized to fill
    the lanes (any size up to 128, several chunks per leaf).  Clustered points
    with ptsMax 128 and maxLevel 4 give leaves of every size, 22 of them beyond
    128 targets (up to 5846); every target must still get its whole near field
    once.  In leaves that large the FP32 coordinates of close pairs put the
    mixed result 6.4e-6 from FP64 (the same with one chunk per 128 targets),
    so the bar is 2e-5: a target missed or counted twice moves it by ~1e-2."""
    n = 30000
    pos = W.gaussians(n, 98)
    den = W.densities(n, 1, 99, kind="normal")
    print("case mixed leaf split n=%d" % n, flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(111, [1.0, 0.25], 4, 128, 4, precision=prec)
        f.set_points(pos)
        f.setup()
        out[prec] = f.evaluate(den)
        out[prec + "_err"] = f.check(den, out[prec], 2000, seed=11)
    d = rel(out["mixed"], out["fp64"])
    print("leaf split: mixed vs fp64 %.3e, errors vs direct: fp64 %.3e mixed %.3e"
          % (d, out["fp64_err"], out["mixed_err"]), flush=True)
    assert d < 2e-5
    assert out["mixed_err"] <= out["fp64_err"] + 2 * d


def test_batched_problems_multi_rhs():
    """The many-problem workload, reduced: independent Navier problems (np 8,
    ptsMax 120, maxLevel 10, nu 0.3), each with 3 right-hand sides through
    kifmm_evaluate_multi, every value vector against the oracle (FP64 bar)."""
    rng = np.random.default_rng(W.seed(4, 0))
    for p, n in enumerate((2048, 3500, 5000)):
        pos = W.uniform(n, W.seed(4, 10 + p))
        nor = np.zeros_like(pos)
        den = rng.standard_normal((3, n * 3))
        print("case batch problem %d n=%d nrhs=3" % (p, n), flush=True)
        f = kifmm.FMM3d(511, [1.0, 0.3], 8, 120, 10)
        f.set_points(pos, nor)
        f.setup()
        vals = f.evaluate_multi(den)
        nue = 8 ** 3 - 6 ** 3
        nuc = 10 ** 3 - 8 ** 3
        pu = f.operator(0).reshape(nuc * 3, nue * 3).T
        pd = f.operator(1).reshape(nue * 3, nue * 3).T
        ops = O.Ops(511, [1.0, 0.3], 8, pu, pd)
        tree = O.Tree(pos, None, pts_max=120, max_level=10)
        for r in range(3):
            ref = O.evaluate(tree, ops, 511, [1.0, 0.3], pos, nor, pos, den[r])
            assert rel(vals[r], ref) <= PARITY


def test_mixed_deterministic():
    """The mixed path is bitwise reproducible: the same densities give the same
    potentials on one handle, on a second handle, and through the transposed
    pass (split-K GEMMs sum their partial products in a fixed order)."""
    n = 30000
    print("case mixed determinism n=%d" % n, flush=True)
    pos = W.gaussians(n, 63)
    den = W.densities(n, 1, 64)
    hs = []
    for _ in range(2):
        f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision="mixed")
        f.set_points(pos)
        f.setup()
        hs.append(f)
    a = hs[0].evaluate(den)
    assert np.array_equal(hs[0].evaluate(den), a)
    assert np.array_equal(hs[1].evaluate(den), a)
    at = hs[0].evaluate_adjoint(den)
    assert np.array_equal(hs[1].evaluate_adjoint(den), at)


@pytest.mark.parametrize("precision,bar", [("fp64", 0.0), ("mixed", 0.0)])
def test_graph_replay(gpu, precision, bar):
    """hipGraph replay (kifmm_set_graph) on a non-default stream: three replays
    with new densities copied in between by eager torch work, each against an
    eager evaluate of the same densities on a second handle.  Bitwise in both
    precisions (the mixed path's split-K partial products are summed in chunk
    order, without atomics)."""
    import torch
    n = 20000
    print("case graph replay %s" % precision, flush=True)
    pos = W.gaussians(n, 61)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision=precision)
    f.set_points(pos)
    f.setup()
    g = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30, precision=precision)
    g.set_points(pos)
    g.setup()
    g.set_graph(True)
    st = torch.cuda.Stream(device=gpu)
    d_den = torch.empty(n, dtype=torch.float64, device=gpu)
    d_val = torch.empty(n, dtype=torch.float64, device=gpu)
    for rep in range(3):
        den = W.densities(n, 1, 62 + rep)
        with torch.cuda.stream(st):
            d_den.copy_(torch.from_numpy(den).to(gpu))
            g.evaluate_device(d_den, d_val, stream=st.cuda_stream)
        st.synchronize()
        got = d_val.cpu().numpy()
        ref = f.evaluate(den)
        r = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print("replay %d relative difference %.3e" % (rep, r), flush=True)
        assert r <= bar


def test_stokes_double_layer_spheres_np6():
    """The sphere-surface workload, reduced: Stokes double layer (kernel 321,
    equivalent kernel 301, sd 4) at np 6 on 32 non-overlapping spheres, ptsMax 64,
    3-vector N(0,1) densities.  FP64 against the oracle at the parity bar; the
    mixed path against the same oracle values (bar 3e-5: the FP32 near field of a
    double layer on a surface, where r.n cancels for close pairs) and no less
    accurate than FP64 against the direct sum beyond that difference."""
    n = 30000
    pos, nor = W.spheres(n, W.seed(1, 0), nsph=32)
    den = W.densities(n, 3, W.seed(1, 1), kind="normal")
    f, val, ref = run_pair(321, [1.0, 0.25], 6, 64, pos, nor, den)
    assert rel(val, ref) <= PARITY
    g = kifmm.FMM3d(321, [1.0, 0.25], 6, 64, 30, precision="mixed")
    g.set_points(pos, nor)
    g.setup()
    mv = g.evaluate(den)
    d = rel(mv, ref)
    e64 = f.check(den, val, 2000, seed=W.seed(1, 3))
    emx = g.check(den, mv, 2000, seed=W.seed(1, 3))
    print("stokes DL np6: mixed vs oracle %.3e, errors vs direct fp64 %.3e mixed %.3e" % (d, e64, emx), flush=True)
    assert d < 3e-5
    assert emx <= e64 + 2 * d


def test_mixed_vs_oracle_clustered_np6():
    """The clustered Laplace workload's settings (16 Gaussians + 10% uniform,
    np 6, ptsMax 128, maxLevel 14) at 40,000 points: the mixed path (FP32 near
    field, S2M/S2L and V list; FP64 pseudo-inverses and sums) against the
    oracle's FP64 restatement of the reference CPU evaluate on the same tree and
    tables.  Bar 1e-6: FP32 arithmetic on the far field; a wrong list, scale or
    table entry moves the result by 1e-3 or more."""
    n = 40000
    pos = W.gaussians(n, W.seed(2, 0))
    den = W.densities(n, 1, W.seed(2, 1))
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 128, 14, precision="mixed")
    f.set_points(pos)
    f.setup()
    val = f.evaluate(den)
    pu = f.operator(0).reshape(296, 152).T
    pd = f.operator(1).reshape(152, 152).T
    ops = O.Ops(111, [1.0, 0.25], 6, pu, pd)
    tree = O.Tree(pos, None, pts_max=128, max_level=14)
    ref = O.evaluate(tree, ops, 111, [1.0, 0.25], pos, np.zeros_like(pos), pos, den)
    assert rel(val, ref) < 1e-6


@pytest.mark.parametrize("kt,coefs", [(111, [1.0, 0.25]), (121, [1.0, 0.25]), (311, [1.0, 0.25]),
                                      (321, [1.0, 0.25]), (322, [1.0, 0.25]), (511, [1.0, 0.3])])
def test_direct_vs_oracle(gpu, kt, coefs):
    """kifmm_direct_device (the FP64 direct sum every accuracy figure is measured
    against, Dense3d_MPI::evaluate, dense3d_mpi.cpp:122-129) against the
    oracle's direct sum at 1,500 targets, sources on sphere surfaces with
    normals; coincident targets exercise the r < 1e-8 exclusion."""
    import torch
    n, m = 12000, 1500
    pos, nor = W.spheres(n, 400 + kt, nsph=8)
    rng = np.random.default_rng(401 + kt)
    tgt = np.concatenate([pos[rng.choice(n, m // 2, replace=False)], W.uniform(m - m // 2, 402 + kt)])
    den = W.densities(n, O.src_dof(kt), 403 + kt, kind="normal")
    print("case direct kt=%d n=%d targets=%d" % (kt, n, m), flush=True)
    f = kifmm.FMM3d(kt, coefs, 4, 64, 30)
    f.set_points(pos, nor)
    f.setup()
    d_den = torch.from_numpy(den).to(gpu)
    d_chk = torch.from_numpy(np.ascontiguousarray(tgt)).to(gpu)
    d_out = torch.empty(m * f.trg_dof, dtype=torch.float64, device=gpu)
    f.direct_device(d_den, d_chk, d_out)
    torch.cuda.synchronize()
    ref = O.direct(kt, coefs, pos, nor, den, tgt)
    assert rel(d_out.cpu().numpy(), ref) <= 1e-13


def test_check_vs_oracle_all_targets():
    """kifmm_check over all targets (numchk >= N, fmm3d_check_mpi.cpp:46-52) is
    the reference's relative 2-norm error (:159-165) against the oracle's
    direct sum, computed here from the same values."""
    n = 6000
    pos = W.uniform(n, 410)
    den = W.densities(n, 1, 411)
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    f.set_points(pos)
    f.setup()
    val = f.evaluate(den)
    err = f.check(den, val, n + 5, seed=1)
    ref = O.direct(111, [1.0, 0.25], pos, np.zeros_like(pos), den, pos)
    want = np.linalg.norm(val - ref) / np.linalg.norm(ref)
    print("check %.6e oracle %.6e" % (err, want), flush=True)
    assert abs(err - want) <= 1e-9 * want


def test_handle_reuse_new_points(gpu):
    """set_points + setup again on one handle with a larger point set: the
    evaluate, the direct sum and the check use the new points (against a fresh
    handle on the same points, bitwise)."""
    import torch
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    a = W.uniform(3000, 420)
    f.set_points(a)
    f.setup()
    f.check(W.densities(3000, 1, 421), f.evaluate(W.densities(3000, 1, 421)), 200, seed=3)
    b = W.gaussians(9000, 422)
    den = W.densities(9000, 1, 423)
    f.set_points(b)
    f.setup()
    g = kifmm.FMM3d(111, [1.0, 0.25], 4, 32, 30)
    g.set_points(b)
    g.setup()
    vf, vg = f.evaluate(den), g.evaluate(den)
    assert np.array_equal(vf, vg)
    assert f.check(den, vf, 500, seed=4) == g.check(den, vg, 500, seed=4)
    d_den = torch.from_numpy(den).to(gpu)
    d_chk = torch.from_numpy(np.ascontiguousarray(b[:700])).to(gpu)
    outs = []
    for h in (f, g):
        o = torch.empty(700, dtype=torch.float64, device=gpu)
        h.direct_device(d_den, d_chk, o)
        outs.append(o)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    ref = O.direct(111, [1.0, 0.25], b, np.zeros_like(b), den, b[:700])
    assert rel(outs[0].cpu().numpy(), ref) <= 1e-13


def test_forest_navier_np8_vs_oracle():
    """The many-problem workload, reduced: four Navier problems (np 8, ptsMax
    120, maxLevel 10, nu 0.3) of uneven size in one handle (one launch per pass
    over all trees), two right-hand sides through kifmm_evaluate_multi; every
    problem's values against the oracle's evaluate of that problem alone, at
    the FP64 bar."""
    sizes = (2048, 5000, 900, 3300)
    probs = [W.uniform(n, W.seed(4, 20 + p)) for p, n in enumerate(sizes)]
    rng = np.random.default_rng(W.seed(4, 1))
    S = sum(sizes)
    den = rng.standard_normal((2, 3 * S))
    print("case forest navier np8 sizes=%s nrhs=2" % (sizes,), flush=True)
    f = kifmm.FMM3d(511, [1.0, 0.3], 8, 120, 10)
    f.set_points_batch(probs)
    f.setup()
    vals = f.evaluate_multi(den)
    nue, nuc = 8 ** 3 - 6 ** 3, 10 ** 3 - 8 ** 3
    pu = f.operator(0).reshape(nuc * 3, nue * 3).T
    pd = f.operator(1).reshape(nue * 3, nue * 3).T
    ops = O.Ops(511, [1.0, 0.3], 8, pu, pd)
    off = 0
    for pos in probs:
        n = len(pos)
        tree = O.Tree(pos, None, pts_max=120, max_level=10)
        for r in range(2):
            ref = O.evaluate(tree, ops, 511, [1.0, 0.3], pos, np.zeros_like(pos), pos, den[r, 3 * off:3 * (off + n)])
            assert rel(vals[r, 3 * off:3 * (off + n)], ref) <= PARITY
        off += n
    # the same forest in mixed precision against the FP64 forest
    g = kifmm.FMM3d(511, [1.0, 0.3], 8, 120, 10, precision="mixed")
    g.set_points_batch(probs)
    g.setup()
    mv = g.evaluate_multi(den)
    assert rel(mv, vals) < 1e-6
