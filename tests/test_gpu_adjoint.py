"""GPU: the transposed evaluate pass (kifmm_evaluate_adjoint) against the forward
operator A of kifmm_evaluate.

The reference has no adjoint, so the pass is pinned two ways:
  * duality: <w, A sigma> = <A^T w, sigma> for random pairs of vectors, with A the
    forward pass that the parity tests pin to the oracle; a wrong transpose, scale,
    list inversion or level order in any step shows as a gap far above round-off;
  * dense: on a small multi-level tree the matrix A is formed column by column from
    the oracle's CPU evaluate (the restatement of fmm3d_eval_mpi.cpp:396-1515), and
    A^T w is compared entry by entry.
"""
import numpy as np
import pytest

import kifmm
import oracle as O
from kifmm import workloads as W

pytestmark = pytest.mark.gpu


def dual_gap(f, seed):
    sd, td = f.src_dof, f.trg_dof
    rng = np.random.default_rng(seed)
    sig = rng.standard_normal(f.n_src * sd)
    w = rng.standard_normal(f.n_trg * td)
    a = f.evaluate(sig)
    at = f.evaluate_adjoint(w)
    scale = np.linalg.norm(w) * np.linalg.norm(a) + np.linalg.norm(at) * np.linalg.norm(sig)
    gap = abs(np.dot(w, a) - np.dot(at, sig)) / scale
    print("duality gap %.3e" % gap, flush=True)
    return gap


ADJ_KERNELS = [(111, [1.0, 0.25], 4), (111, [1.0, 0.25], 6), (121, [1.0, 0.25], 4), (311, [1.0, 0.25], 4),
               (312, [1.0, 0.25], 4), (321, [1.0, 0.25], 4), (321, [1.0, 0.25], 6), (322, [1.0, 0.25], 4),
               (511, [1.0, 0.3], 4), (521, [1.0, 0.3], 4)]


@pytest.mark.parametrize("kt,coefs,np_", ADJ_KERNELS)
def test_adjoint_duality_fp64(kt, coefs, np_):
    """Adaptive sphere-surface points (U, V, W and X lists all active)."""
    n = 6000
    print("case adjoint kt=%d np=%d n=%d" % (kt, np_, n), flush=True)
    pos, nor = W.spheres(n, 31 + kt, nsph=6)
    f = kifmm.FMM3d(kt, coefs, np_, 24, 30)
    f.set_points(pos, nor)
    f.setup()
    bar = 1e-12 if np_ == 4 else 1e-10  # np 6: the pseudo-inverse amplifies round-off
    for seed in (1, 2):
        assert dual_gap(f, seed) < bar


def test_adjoint_duality_separate_targets():
    """Sources and targets differ (the inverted lists are not the forward ones)."""
    print("case adjoint separate targets", flush=True)
    src = W.uniform(5000, 41)
    trg = W.gaussians(4000, 42)
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 20, 30)
    f.set_points(src, None, trg)
    f.setup()
    assert dual_gap(f, 3) < 1e-12


@pytest.mark.parametrize("kt,coefs", [(111, [1.0, 0.25]), (511, [1.0, 0.3]), (321, [1.0, 0.25])])
def test_adjoint_duality_mixed(kt, coefs):
    """Mixed precision: FP32 forward pieces bound the gap."""
    n = 20000
    print("case adjoint mixed kt=%d" % kt, flush=True)
    if kt == 321:
        pos, nor = W.spheres(n, 43, nsph=5)
    else:
        pos, nor = W.gaussians(n, 43), None
    f = kifmm.FMM3d(kt, coefs, 6, 32, 30, precision="mixed")
    f.set_points(pos, nor)
    f.setup()
    assert dual_gap(f, 4) < 1e-5


@pytest.mark.parametrize("case", ["gaussians", "spheres", "separate"])
def test_adjoint_mixed_vs_fp64(case):
    """Laplace mixed transposed pass (FP32 near field through the forward near
    kernel with the roles of sources and targets swapped) against the FP64
    transposed pass on the same tree: U, W and X lists active (sphere surfaces),
    and sources distinct from targets (the swap must read target points for the
    pairs and source points for the chunks).  Bar: the mixed forward pass's own
    difference from FP64 on the same points (the FP32 near field sets both; on
    sphere surfaces it is 2e-6, on Gaussian clusters 1e-7)."""
    print("case adjoint mixed vs fp64 %s" % case, flush=True)
    src, trg, nor = None, None, None
    if case == "gaussians":
        src = W.gaussians(20000, 61)
    elif case == "spheres":
        src, nor = W.spheres(20000, 62, nsph=6)
    else:
        src, trg = W.uniform(12000, 63), W.gaussians(9000, 64)
    fwd, adj = {}, {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(111, [1.0, 0.25], 6, 32, 30, precision=prec)
        f.set_points(src, nor, trg)
        f.setup()
        rng = np.random.default_rng(65)
        fwd[prec] = f.evaluate(rng.standard_normal(f.n_src))
        adj[prec] = f.evaluate_adjoint(rng.standard_normal(f.n_trg))
        f.close()
    rf = np.linalg.norm(fwd["mixed"] - fwd["fp64"]) / np.linalg.norm(fwd["fp64"])
    ra = np.linalg.norm(adj["mixed"] - adj["fp64"]) / np.linalg.norm(adj["fp64"])
    print("relative difference: forward %.3e, transposed %.3e" % (rf, ra), flush=True)
    assert ra < 1.5 * rf + 1e-7 and ra < 1e-5


@pytest.mark.parametrize("case", ["gaussians", "spheres"])
def test_field_mixed_vs_fp64(case):
    """Laplace mixed field pass against the FP64 field pass on the same tree.
    The mixed field keeps its near field in FP64 (an FP32 gradient near field
    lost two orders of accuracy on sphere surfaces, where close pairs dominate
    the 1/r^2 terms), so only its far field is mixed: well inside the mixed
    forward pass's own difference from FP64."""
    print("case field mixed vs fp64 %s" % case, flush=True)
    nor = None
    if case == "gaussians":
        src = W.gaussians(20000, 66)
    else:
        src, nor = W.spheres(20000, 67, nsph=6)
    fwd, fld = {}, {}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(111, [1.0, 0.25], 6, 32, 30, precision=prec)
        f.set_points(src, nor)
        f.setup()
        den = W.densities(f.n_src, 1, 68)
        fwd[prec] = f.evaluate(den)
        fld[prec] = f.field(den)
        f.close()
    rf = np.linalg.norm(fwd["mixed"] - fwd["fp64"]) / np.linalg.norm(fwd["fp64"])
    rg = np.linalg.norm(fld["mixed"] - fld["fp64"]) / np.linalg.norm(fld["fp64"])
    print("relative difference: forward %.3e, field %.3e" % (rf, rg), flush=True)
    assert rg < 1e-6 and rg < rf


@pytest.mark.parametrize("np_,bar", [(4, 1e-12), (6, 1e-10)])
def test_adjoint_dense_vs_oracle(np_, bar):
    """A formed from the oracle's evaluate, one unit density at a time.  At np 6
    the bar is the duality bar: the pseudo-inverses (condition ~1e7) amplify the
    round-off of the transposed summation order."""
    n, pm = 240, 6
    print("case adjoint dense n=%d np=%d" % (n, np_), flush=True)
    pos = W.gaussians(n, 44)
    nor = np.zeros_like(pos)
    f = kifmm.FMM3d(111, [1.0, 0.25], np_, pm, 30)
    f.set_points(pos, nor)
    f.setup()
    nue = np_ ** 3 - (np_ - 2) ** 3
    nuc = (np_ + 2) ** 3 - np_ ** 3
    pu = f.operator(0).reshape(nuc, nue).T
    pd = f.operator(1).reshape(nue, nue).T
    ops = O.Ops(111, [1.0, 0.25], np_, pu, pd)
    tree = O.Tree(pos, None, pts_max=pm, max_level=30)
    assert tree.lists["v"][1].size > 0, "the tree must have V lists"
    A = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = O.evaluate(tree, ops, 111, [1.0, 0.25], pos, nor, pos, e)
    w = np.random.default_rng(5).standard_normal(n)
    at = f.evaluate_adjoint(w)
    ref = A.T @ w
    r = np.linalg.norm(at - ref) / np.linalg.norm(ref)
    print("relative difference %.3e" % r, flush=True)
    assert r < bar


def direct_grad(pos, den, tgt):
    """FP64 gradient of sum_s den_s / (4 pi |t - s|) at the rows of tgt (pairs
    closer than 1e-8 excluded, as the reference kernel does)."""
    d = tgt[:, None, :] - pos[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    inv3 = np.where(r2 < 1e-16, 0.0, np.maximum(r2, 1e-16) ** -1.5)
    return -np.einsum("ij,j,ijk->ik", inv3, den, d) / (4 * np.pi)


@pytest.mark.parametrize("np_,precision,bar", [(6, "fp64", 3e-4), (8, "fp64", 3e-5), (6, "mixed", 3e-4)])
def test_field_vs_direct(np_, precision, bar):
    """Gradient pass (kifmm_field) against the FP64 direct gradient at 300 targets."""
    n = 20000
    print("case field np=%d %s" % (np_, precision), flush=True)
    pos = W.gaussians(n, 45)
    den = W.densities(n, 1, 46)
    f = kifmm.FMM3d(111, [1.0, 0.25], np_, 40, 30, precision=precision)
    f.set_points(pos)
    f.setup()
    g = f.field(den)
    idx = np.random.default_rng(7).choice(n, 300, replace=False)
    ref = direct_grad(pos, den, pos[idx])
    r = np.linalg.norm(g[idx] - ref) / np.linalg.norm(ref)
    print("relative error %.3e" % r, flush=True)
    assert r < bar


def test_field_rejects_other_kernels():
    pos, nor = W.spheres(500, 47, nsph=2)
    f = kifmm.FMM3d(321, [1.0, 0.25], 4, 32, 30)
    f.set_points(pos, nor)
    f.setup()
    with pytest.raises(kifmm.KifmmError) as e:
        f.field(np.zeros(500 * 3))
    assert e.value.code == -2


def test_autograd_gradcheck():
    """torch.autograd through fmm_apply: backward (the transposed pass) against
    finite differences of the forward pass (exact for a linear map up to
    rounding; eps = 1e-4 keeps the rounding of O(10) potentials, ~1e-14, far
    below the tolerance after division by 2 eps)."""
    import torch
    from kifmm.autograd import fmm_apply
    n = 300
    print("case autograd n=%d" % n, flush=True)
    pos = W.gaussians(n, 48)
    f = kifmm.FMM3d(111, [1.0, 0.25], 4, 8, 30)
    f.set_points(pos)
    f.setup()
    den = torch.tensor(W.densities(n, 1, 49), dtype=torch.float64, device="cuda:0", requires_grad=True)
    assert torch.autograd.gradcheck(lambda d: fmm_apply(f, d), (den,), eps=1e-4, atol=1e-7, rtol=1e-6)


def test_autograd_loss_gradient():
    """The autograd wiring: for L = 1/2 |A sigma - A tau|^2, backward hands
    grad_phi = phi - psi to the transposed pass and returns its result, so
    dL/dsigma is bitwise the transposed pass of phi - psi (the FP64 passes are
    deterministic).  The transposed pass itself is pinned by the duality and
    dense-oracle tests above and the finite differences below."""
    import torch
    from kifmm.autograd import fmm_apply
    n = 20000
    print("case autograd loss n=%d" % n, flush=True)
    pos = W.uniform(n, 50)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30)
    f.set_points(pos)
    f.setup()
    dev = torch.device("cuda:0")
    sig = torch.tensor(W.densities(n, 1, 51), dtype=torch.float64, device=dev, requires_grad=True)
    tau = torch.tensor(W.densities(n, 1, 52), dtype=torch.float64, device=dev)
    phi = fmm_apply(f, sig)
    psi = fmm_apply(f, tau)
    loss = 0.5 * torch.sum((phi - psi) ** 2)
    loss.backward()
    ref = f.evaluate_adjoint((phi - psi).detach().cpu().numpy())
    assert np.array_equal(sig.grad.cpu().numpy(), ref)


def test_position_grad_vs_direct():
    """Position gradient of L = 1/2 |A sigma - A tau|^2 from two field passes
    (kifmm.autograd.position_grad) against the same closed form on direct sums."""
    from kifmm.autograd import position_grad
    from test_gradient_formula import apply, field
    n = 3000
    print("case position gradient n=%d" % n, flush=True)
    x = W.uniform(n, W.seed(3, 0))
    sig = W.densities(n, 1, W.seed(3, 1))
    tau = W.densities(n, 1, W.seed(3, 2))
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 32, 30)
    f.set_points(x)
    f.setup()
    u = sig - tau
    pg = position_grad(f, u, f.evaluate(u))
    g = apply(x, u)
    ref = g[:, None] * field(x, u) + u[:, None] * field(x, g)
    r = np.linalg.norm(pg - ref) / np.linalg.norm(ref)
    print("relative error %.3e" % r, flush=True)
    assert r < 1e-3


def test_adjoint_dense_vs_oracle_stokes():
    """Stokes double layer (kernel 321, equivalent kernel 301 with its odd
    pressure-doublet column): A formed from the oracle's evaluate one unit
    density at a time on a two-sphere surface, A^T w against the transposed
    pass (np 4)."""
    n, np_, pm = 150, 4, 6
    print("case adjoint dense stokes n=%d" % n, flush=True)
    pos, nor = W.spheres(n, 53, nsph=2)
    f = kifmm.FMM3d(321, [1.0, 0.25], np_, pm, 30)
    f.set_points(pos, nor)
    f.setup()
    nue = np_ ** 3 - (np_ - 2) ** 3
    nuc = (np_ + 2) ** 3 - np_ ** 3
    pu = f.operator(0).reshape(nuc * 3, nue * 4).T
    pd = f.operator(1).reshape(nue * 3, nue * 4).T
    ops = O.Ops(301, [1.0, 0.25], np_, pu, pd)
    tree = O.Tree(pos, None, pts_max=pm, max_level=30)
    assert tree.lists["v"][1].size > 0, "the tree must have V lists"
    A = np.empty((3 * n, 3 * n))
    for j in range(3 * n):
        e = np.zeros(3 * n)
        e[j] = 1.0
        A[:, j] = O.evaluate(tree, ops, 321, [1.0, 0.25], pos, nor, pos, e)
    w = np.random.default_rng(54).standard_normal(3 * n)
    at = f.evaluate_adjoint(w)
    ref = A.T @ w
    r = np.linalg.norm(at - ref) / np.linalg.norm(ref)
    print("relative difference %.3e" % r, flush=True)
    assert r < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_field_from_checkpoint(gpu, precision):
    """The field pass from the forward pass's checkpoint of the upward
    equivalent densities equals the field pass that runs the up pass again
    (bitwise: the same UE), and the checkpointed forward values equal a plain
    evaluate."""
    import torch
    n = 15000
    print("case checkpoint %s" % precision, flush=True)
    pos = W.gaussians(n, 55)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 40, 30, precision=precision)
    f.set_points(pos)
    f.setup()
    d = torch.tensor(W.densities(n, 1, 56), dtype=torch.float64, device=gpu)
    v = torch.empty(n, dtype=torch.float64, device=gpu)
    ue = torch.empty(f.checkpoint_size(), dtype=torch.float64, device=gpu)
    f.evaluate_checkpoint_device(d, v, ue)
    g1 = torch.empty(3 * n, dtype=torch.float64, device=gpu)
    g2 = torch.empty(3 * n, dtype=torch.float64, device=gpu)
    f.field_checkpoint_device(d, ue, g1)
    f.field_device(d, g2)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
    assert np.array_equal(v.cpu().numpy(), f.evaluate(d.cpu().numpy()))


def test_autograd_positions_small(gpu):
    """fmm_apply with positions: x.grad of L = 1/2 |A(x) sigma - A(x) tau|^2
    equals the two-field closed form g grad(A u) + u grad(A g) (u = sigma - tau,
    g = A u) evaluated separately; the autograd form runs four field passes
    (two from checkpoints) whose sum is that one by linearity."""
    import torch
    from kifmm.autograd import fmm_apply, position_grad
    n = 8000
    print("case autograd positions n=%d" % n, flush=True)
    pos = W.uniform(n, 57)
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 40, 30)
    f.set_points(pos)
    f.setup()
    sig = W.densities(n, 1, 58)
    tau = W.densities(n, 1, 59)
    x = torch.tensor(pos, dtype=torch.float64, device=gpu, requires_grad=True)
    s = torch.tensor(sig, dtype=torch.float64, device=gpu, requires_grad=True)
    t = torch.tensor(tau, dtype=torch.float64, device=gpu)
    r = fmm_apply(f, s, x) - fmm_apply(f, t, x)
    loss = 0.5 * torch.sum(r * r)
    loss.backward()
    u = sig - tau
    ref = position_grad(f, u, f.evaluate(u))
    d = np.linalg.norm(x.grad.cpu().numpy() - ref) / np.linalg.norm(ref)
    print("relative difference %.3e" % d, flush=True)
    assert d < 1e-9


def test_gradients_full_size_finite_differences(gpu):
    """The gradient workload at full size: N = 1,048,576 uniform points, np 6,
    L = 1/2 sum_i (phi_i - psi_i)^2 with phi = A(x) sigma, psi = A(x) tau through
    fmm_apply (FP64).  Central differences with h = 1e-6 (tree and leaves
    frozen; self pairs excluded by index) on 32 density and 32 position
    coordinates, summing only the terms of L that change:
      density k: phi moves by h A e_k (A e_k from an FMM evaluate), so the
        difference is exact up to rounding (bar 1e-7);
      position (k, d): row k and column k of the direct-sum operator move,
        r_i += u_k (G(x_i - x_k') - G(x_i - x_k)) (i != k) and
        r_k += sum_j u_j (G(x_k' - x_j) - G(x_k - x_j)) (j != k), u = sigma - tau,
        against the FMM field passes (bar: 2e-4 of the largest component,
        the np-6 gradient accuracy)."""
    import torch
    from kifmm.autograd import fmm_apply
    n = 1048576
    print("case full-size gradients n=%d" % n, flush=True)
    pos = W.uniform(n, W.seed(3, 0))
    sig = W.densities(n, 1, W.seed(3, 1))
    tau = W.densities(n, 1, W.seed(3, 2))
    f = kifmm.FMM3d(111, [1.0, 0.25], 6, 64, 30)
    f.set_points(pos)
    f.setup()
    x = torch.tensor(pos, dtype=torch.float64, device=gpu, requires_grad=True)
    s = torch.tensor(sig, dtype=torch.float64, device=gpu, requires_grad=True)
    t = torch.tensor(tau, dtype=torch.float64, device=gpu)
    r_t = fmm_apply(f, s, x) - fmm_apply(f, t, x)
    loss = 0.5 * torch.sum(r_t * r_t)
    loss.backward()
    r = r_t.detach().cpu().numpy()
    gs = s.grad.cpu().numpy()
    gx = x.grad.cpu().numpy()
    rng = np.random.default_rng(W.seed(3, 4))
    h = 1e-6
    # densities
    ks = rng.choice(n, 32, replace=False)
    e = torch.zeros(n, dtype=torch.float64, device=gpu)
    c = torch.empty(n, dtype=torch.float64, device=gpu)
    worst = 0.0
    for k in ks:
        e.zero_()
        e[int(k)] = 1.0
        f.evaluate_device(e, c)
        cv = c.cpu().numpy()
        fd = np.sum((r + h * cv) ** 2 - (r - h * cv) ** 2) / (4 * h)
        worst = max(worst, abs(fd - gs[k]) / max(abs(fd), 1e-3 * np.abs(gs).max()))
    print("density gradient: worst relative difference %.3e" % worst, flush=True)
    assert worst < 1e-7
    # positions
    u = sig - tau
    four_pi = 4.0 * np.pi

    def dL(k, xk_new):
        dxi = pos - xk_new
        dxo = pos - pos[k]
        ri = np.sqrt(np.einsum("ij,ij->i", dxi, dxi))
        ro = np.sqrt(np.einsum("ij,ij->i", dxo, dxo))
        ri[k] = ro[k] = 1.0
        dg = (1.0 / ri - 1.0 / ro) / four_pi
        dg[k] = 0.0
        delta = u[k] * dg            # column k: every other target
        delta[k] = np.dot(u, dg)     # row k: the moved target (the kernel is symmetric)
        return np.sum(delta * (r + 0.5 * delta))

    pts = rng.choice(n, 32, replace=False)
    axes = rng.integers(0, 3, 32)
    fds, ags = [], []
    for k, d in zip(pts, axes):
        xp, xm = pos[k].copy(), pos[k].copy()
        xp[d] += h
        xm[d] -= h
        fds.append((dL(k, xp) - dL(k, xm)) / (2 * h))
        ags.append(gx[k, d])
    fds, ags = np.array(fds), np.array(ags)
    err = np.abs(fds - ags).max() / np.abs(fds).max()
    print("position gradient: max difference %.3e of the largest component" % err, flush=True)
    assert err < 2e-4
