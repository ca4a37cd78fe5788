"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors (its only check is the direct sum,
fmm3d_check_mpi.cpp), so each oracle function is pinned by mathematics that a
plausible mistake (dropped term, wrong sign, transposed operand, wrong index)
would break:
  * FFT helpers     -> numpy.fft.rfftn / irfftn (library routines);
  * kernels         -> PDE / integral identities and homogeneity (closed forms);
  * geometry        -> counts and boundary membership;
  * FFT M2L         -> the dense kernel product on the same surfaces;
  * tree and lists  -> brute-force KIFMM completeness and symmetry invariants;
  * evaluate        -> the direct sum (np-dependent accuracy, exact for N <= ptsMax).
"""
import itertools

import numpy as np
import pytest

import oracle as O
from kifmm import workloads as W

rng = np.random.default_rng(1234)


# ---------------------------------------------------------------- FFT -------
@pytest.mark.parametrize("n,howmany", [(8, 1), (12, 4), (16, 3)])
def test_rfft3_matches_numpy(n, howmany):
    x = rng.standard_normal(n ** 3 * howmany)
    X = O.rfft3(x, n, howmany).view(np.complex128).reshape(n, n, n // 2 + 1, howmany)
    ref = np.fft.rfftn(x.reshape(n, n, n, howmany), axes=(0, 1, 2))
    assert np.abs(X - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n,howmany", [(8, 1), (12, 3)])
def test_irfft3_matches_numpy(n, howmany):
    x = rng.standard_normal((n, n, n, howmany))
    X = np.fft.rfftn(x, axes=(0, 1, 2))
    back = O.irfft3(X.reshape(-1).view(np.float64), n, howmany).reshape(n, n, n, howmany)
    assert np.abs(back / n ** 3 - x).max() < 1e-13  # unnormalised like FFTW's c2r


# ------------------------------------------------------------- kernels ------
def kmat(kt, coefs, src, nor, trg):
    return O.kernel_matrix(kt, coefs, np.atleast_2d(src), np.atleast_2d(nor), np.atleast_2d(trg))


def sphere_quadrature(nth=40, nph=80, radius=1.0, centre=(0, 0, 0)):
    """Gauss-Legendre in cos(theta) x trapezoid in phi: spectral for smooth integrands."""
    u, wu = np.polynomial.legendre.leggauss(nth)
    ph = 2 * np.pi * np.arange(nph) / nph
    U, PH = np.meshgrid(u, ph, indexing="ij")
    st = np.sqrt(1 - U ** 2)
    nrm = np.stack([st * np.cos(PH), st * np.sin(PH), U], -1).reshape(-1, 3)
    w = (np.repeat(wu, nph) * (2 * np.pi / nph)) * radius ** 2
    return np.asarray(centre) + radius * nrm, nrm, w


def double_layer_integral(kt, coefs, x):
    pts, nrm, w = sphere_quadrature()
    sd, td = O.src_dof(kt), O.trg_dof(kt)
    M = kmat(kt, coefs, pts, nrm, x).reshape(td, len(pts), sd)
    return np.einsum("tqs,q->ts", M, w)


def test_laplace_double_layer_gauss_law():
    # 121: -(1/4pi) (r.n)/r^3, r = target - source: unit density gives 1 inside, 0 outside
    assert abs(double_layer_integral(121, [1, .25], [0.2, -0.1, 0.3])[0, 0] - 1.0) < 1e-10
    assert abs(double_layer_integral(121, [1, .25], [1.7, 0.4, -0.2])[0, 0]) < 1e-10


def test_stokes_double_layer_identity():
    # 321: integral of the stresslet against outward normals is I inside, 0 outside
    np.testing.assert_allclose(double_layer_integral(321, [1, .25], [0.1, 0.2, -0.3]), np.eye(3), atol=1e-9)
    np.testing.assert_allclose(double_layer_integral(321, [1, .25], [0.0, 2.0, 0.5]), 0, atol=1e-9)


def test_navier_double_layer_rigid_translation():
    # 521 (Kelvin traction kernel): a rigid translation is reproduced exactly inside
    # (Somigliana identity with zero traction), nothing outside
    for nu in (0.25, 0.3):
        I_in = double_layer_integral(521, [1, nu], [0.15, -0.2, 0.1])
        I_out = double_layer_integral(521, [1, nu], [1.9, 0.0, 0.3])
        np.testing.assert_allclose(np.abs(I_in), np.eye(3), atol=1e-9)
        np.testing.assert_allclose(I_out, 0, atol=1e-9)
        assert abs(abs(np.trace(I_in)) - 3) < 1e-9 and np.allclose(I_in, I_in[0, 0] * np.eye(3), atol=1e-9)


def fd_field(kt, coefs, x0, src=np.zeros(3), nor=np.array([0.3, -0.5, 0.8]), h=1e-3):
    """Kernel block and its first and second derivatives in the target by central differences."""
    def K(x):
        return kmat(kt, coefs, src, nor, x)
    d1, d2 = [], []
    for a in range(3):
        e = np.zeros(3); e[a] = h
        d1.append((K(x0 + e) - K(x0 - e)) / (2 * h))
        d2.append((K(x0 + e) - 2 * K(x0) + K(x0 - e)) / h ** 2)
    return K(x0), d1, d2


@pytest.mark.parametrize("kt", [111, 121, 312, 322])
def test_scalar_kernels_are_harmonic(kt):
    x0 = np.array([0.37, -0.21, 0.44])
    _, _, d2 = fd_field(kt, [1.0, 0.25], x0)
    lap = sum(d2)
    scale = max(np.abs(d).max() for d in d2)   # a non-harmonic formula gives |lap| ~ scale
    assert np.abs(lap).max() < 1e-4 * scale


def test_stokeslet_solves_stokes_with_its_pressure():
    # mu Lap u - grad p = 0 and div u = 0 away from the source (311 velocity, 312 pressure)
    mu = 1.7
    x0 = np.array([0.31, 0.27, -0.4])
    _, d1, d2 = fd_field(311, [mu, 0.25], x0)
    lap_u = sum(d2)                     # 3x3: d^2 u_i / dx^2 for force direction j
    _, dp, _ = fd_field(312, [mu, 0.25], x0)
    grad_p = np.stack([dp[i][0] for i in range(3)])  # (i, j)
    scale2 = max(np.abs(mu * lap_u).max(), np.abs(grad_p).max(), max(np.abs(mu * d).max() for d in d2))
    assert np.abs(mu * lap_u - grad_p).max() < 1e-4 * scale2
    div = sum(d1[i][i] for i in range(3))  # sum_i d u_i/dx_i for each j
    assert np.abs(div).max() < 1e-5 * max(np.abs(d).max() for d in d1)


def test_stokes_fmm_kernel_blocks():
    # 301 = [Stokeslet(mu) | r/(4 pi r^3)]: velocity part equals 311, last column div-free
    x0 = np.array([0.2, -0.35, 0.3])
    for mu in (1.0, 2.5):
        K301 = kmat(301, [mu, 0.25], np.zeros(3), np.zeros(3), x0)
        K311 = kmat(311, [mu, 0.25], np.zeros(3), np.zeros(3), x0)
        np.testing.assert_allclose(K301[:, :3], K311, rtol=1e-14)
    _, d1, _ = fd_field(301, [1.0, 0.25], x0)
    scale = max(abs(d1[i][i, 3]) for i in range(3))
    assert abs(sum(d1[i][i, 3] for i in range(3))) < 1e-4 * scale


@pytest.mark.parametrize("nu", [0.25, 0.3])
def test_kelvin_solution_solves_navier(nu):
    # mu Lap u + (lambda + mu) grad div u = 0, lambda = 2 mu nu / (1 - 2 nu)
    mu = 1.3
    lam = 2 * mu * nu / (1 - 2 * nu)
    x0 = np.array([0.29, -0.33, 0.21])
    h = 1e-3
    def K(x):
        return kmat(511, [mu, nu], np.zeros(3), np.zeros(3), x)
    lap = sum((K(x0 + h * e) - 2 * K(x0) + K(x0 - h * e)) / h ** 2 for e in np.eye(3))
    # grad(div u): second mixed derivatives
    gd = np.zeros((3, 3))
    for i in range(3):
        for a in range(3):
            ei, ea = np.eye(3)[i] * h, np.eye(3)[a] * h
            dd = (K(x0 + ei + ea) - K(x0 + ei - ea) - K(x0 - ei + ea) + K(x0 - ei - ea)) / (4 * h * h)
            gd[i] += dd[a]
    scale = max(np.abs(mu * lap).max(), np.abs((lam + mu) * gd).max())
    assert np.abs(mu * lap + (lam + mu) * gd).max() < 1e-4 * scale


@pytest.mark.parametrize("kt,deg", [(111, [1]), (301, [1, 1, 1, 2]), (511, [1, 1, 1])])
def test_homogeneity(kt, deg):
    # K(2r) = 2^-deg K(r) per source DOF (kernel3d_mpi.cpp:122-131)
    x = np.array([0.3, -0.2, 0.5])
    K1 = kmat(kt, [1.0, 0.3], np.zeros(3), np.zeros(3), x)
    K2 = kmat(kt, [1.0, 0.3], np.zeros(3), np.zeros(3), 2 * x)
    np.testing.assert_allclose(K2, K1 * 2.0 ** -np.array(deg)[None, :], rtol=1e-14)


@pytest.mark.parametrize("kt", [111, 311, 511])
def test_single_layer_symmetry(kt):
    a, b = np.array([0.1, 0.2, 0.3]), np.array([0.6, -0.1, 0.2])
    np.testing.assert_allclose(kmat(kt, [1, .3], a, a, b), kmat(kt, [1, .3], b, b, a).T, rtol=1e-14)


def test_singular_pairs_excluded():
    for kt in (111, 121, 301, 311, 312, 321, 322, 511, 521):
        K = kmat(kt, [1, .3], np.full(3, 0.4), np.array([0, 0, 1.0]), np.full(3, 0.4) + 1e-9)
        assert np.all(K == 0), kt


# ------------------------------------------------------------- geometry -----
@pytest.mark.parametrize("np_", [4, 6, 8])
def test_surface_samples(np_):
    for R in (1.0, 3.0):
        p = O.sample_pos(np_, R)
        assert len(p) == np_ ** 3 - (np_ - 2) ** 3
        assert np.allclose(np.abs(p).max(axis=1), R)   # every sample on the cube surface
        assert len({tuple(r) for r in np.round(p, 12)}) == len(p)
    g = O.regular_pos(np_, 1.0)
    assert len(g) == (2 * np_) ** 3 and np.isclose(g.max(), 2.0) and np.isclose(g.min(), -2.0 * np_ / (np_ - 1))


# ------------------------------------------------------------------ M2L -----
@pytest.mark.parametrize("kt,np_", [(111, 4), (111, 6), (301, 4), (511, 4)])
def test_fft_m2l_equals_dense_product(kt, np_):
    """DC(B) from UE(A) through the grid equals the dense K(DC points of B, UE points of A)."""
    coefs = [1.0, 0.3]
    up, dn = O.numpy_pinv_tables(kt, coefs, np_)
    ops = O.Ops(kt, coefs, np_, up, dn)
    level = 3
    r = 2.0 ** -level
    ue = rng.standard_normal(ops.n_ue * ops.sd)
    deg = {111: [1], 301: [1, 1, 1, 2], 511: [1, 1, 1]}[kt]
    for idx in [(2, 0, 0), (-3, 1, 2), (0, -2, 3), (3, 3, -3)]:
        dc = O.m2l_fft(ops, idx, level, ue)
        src = np.array(idx) * 2 * r + r * O.sample_pos(np_, 1.0)      # UE of A (radius r)
        trg = r * O.sample_pos(np_, 1.0)                               # DC of B (radius r)
        ref = kmat(kt, coefs, src, src, trg) @ ue
        np.testing.assert_allclose(dc, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
        assert deg  # scaling exercised through level != 0


# ------------------------------------------------------------ tree / lists --
def leaves(t):
    return [i for i in range(t.n_nodes) if t.child[i] < 0]


def ancestors_or_self(t, i):
    out = []
    while i >= 0:
        out.append(i)
        i = t.parent[i]
    return out


def cells(pos, L):
    return np.minimum((pos * 2 ** L).astype(np.int64), 2 ** L - 1)


@pytest.mark.parametrize("gen,pm,ml", [("uniform", 16, 30), ("gauss", 12, 30), ("gauss", 12, 5)])
def test_tree_invariants(gen, pm, ml):
    n = 3000
    pos = W.uniform(n, 3) if gen == "uniform" else W.gaussians(n, 4)
    t = O.Tree(pos, pts_max=pm, max_level=ml)
    N = t.n_nodes
    # structure: root 0, children blocks of 8 after their parent, child index 4x+2y+z
    assert t.parent[0] == -1 and t.child[0] == 1
    for i in range(N):
        if t.child[i] >= 0:
            c0 = t.child[i]
            assert c0 > i
            for c in range(8):
                assert t.parent[c0 + c] == i and t.depth[c0 + c] == t.depth[i] + 1
                assert tuple(t.path[c0 + c]) == tuple(2 * t.path[i] + [(c >> 2) & 1, (c >> 1) & 1, c & 1])
    # Dendro rule on distinct depth-maxLevel cells: leaves hold <= ptsMax of them
    # unless at maxLevel; split nodes (other than the root) hold more
    cl = cells(pos, ml)
    for i in range(1, N):
        d, p = t.depth[i], t.path[i]
        inside = np.all((cl >> (ml - d)) == p, axis=1)
        distinct = len({tuple(c) for c in cl[inside]})
        if t.child[i] < 0:
            assert distinct <= pm or d == ml
        else:
            assert distinct > pm and d < ml
    # every point lies inside its leaf and leaves are packed in node order
    for i in leaves(t):
        if t.src_num[i] > 0:
            idx = t.src_perm[t.src_beg[i]:t.src_beg[i] + t.src_num[i]]
            lo = t.path[i] * 0.5 ** t.depth[i]
            assert np.all(pos[idx] >= lo) and np.all(pos[idx] < lo + 0.5 ** t.depth[i])
            assert np.all(np.diff(idx) > 0)
    assert sorted(t.src_perm.tolist()) == list(range(n))


@pytest.mark.parametrize("gen,pm", [("uniform", 8), ("gauss", 6), ("spheres", 8)])
def test_lists_cover_every_pair_once(gen, pm):
    """KIFMM completeness: each (target leaf B, source leaf A) interaction is
    evaluated exactly once through U(B), W(B) (A or an ancestor of A), V of an
    ancestor-or-self pair, or X of an ancestor-or-self of B."""
    n = 600
    pos = {"uniform": lambda: W.uniform(n, 8), "gauss": lambda: W.gaussians(n, 9),
           "spheres": lambda: W.spheres(n, 10, nsph=4)[0]}[gen]()
    t = O.Tree(pos, pts_max=pm)
    L = [i for i in leaves(t) if t.trg_num[i] > 0]
    S = [i for i in leaves(t) if t.src_num[i] > 0]
    sets = {k: [set(t.list_of(k, i).tolist()) for i in range(t.n_nodes)] for k in "uvwx"}
    for B in L:
        ancB = ancestors_or_self(t, B)
        for A in S:
            ancA = ancestors_or_self(t, A)
            cnt = (A in sets["u"][B]) + sum(a in sets["w"][B] for a in ancA)
            cnt += sum(a in sets["v"][b] for b in ancB for a in ancA)
            cnt += sum(A in sets["x"][b] for b in ancB)
            assert cnt == 1, (B, A, cnt)
    # symmetry / duality with sources == targets
    for B in range(t.n_nodes):
        for A in sets["u"][B]:
            assert B in sets["u"][A]
        for A in sets["v"][B]:
            assert B in sets["v"][A] and t.depth[A] == t.depth[B]
        for A in sets["w"][B]:
            assert B in sets["x"][A]
        assert len(sets["v"][B]) <= 189
        for k in "uvwx":
            lst = t.list_of(k, B)
            assert np.all(np.diff(lst) > 0)   # ascending node index (std::set order)


# -------------------------------------------------------------- evaluate ----
@pytest.mark.parametrize("kt,np_,coefs,tol", [(111, 4, [1, .25], 5e-4), (111, 6, [1, .25], 1e-5),
                                               (321, 4, [1, .25], 5e-3), (511, 6, [1, .3], 1e-4)])
def test_oracle_fmm_vs_direct(kt, np_, coefs, tol):
    n = 3000
    if kt == 321:
        pos, nor = W.spheres(n, 21, nsph=5)
    else:
        pos, nor = W.uniform(n, 22), np.zeros((n, 3))
    den = W.densities(n, O.src_dof(kt), 23, kind="normal")
    km = O.mm_kernel(kt)
    up, dn = O.numpy_pinv_tables(km, coefs, np_)
    v = O.evaluate(O.Tree(pos, pts_max=32), O.Ops(km, coefs, np_, up, dn), kt, coefs, pos, nor, pos, den)
    d = O.direct(kt, coefs, pos, nor, den, pos)
    assert np.linalg.norm(v - d) / np.linalg.norm(d) < tol


def test_oracle_known_answer_single_level():
    n = 40
    pos = W.uniform(n, 31)
    den = W.densities(n, 1, 32)
    up, dn = O.numpy_pinv_tables(111, [1, .25], 4)
    v = O.evaluate(O.Tree(pos, pts_max=64), O.Ops(111, [1, .25], 4, up, dn), 111, [1, .25], pos,
                   np.zeros_like(pos), pos, den)
    d = O.direct(111, [1, .25], pos, np.zeros_like(pos), den, pos)
    assert np.linalg.norm(v - d) / np.linalg.norm(d) < 1e-14


def check_truncated_pinv(M, P, eps):
    """Truncated Moore-Penrose identities in the spectral norm: M P M differs
    from M by the largest dropped singular value (<= eps*|M|) plus the round-off
    of singular vectors near the cutoff (LAPACK itself reaches 3.4 eps*|M| at
    np 8), P M P = P, P M P = P up to round-off times the condition number, and |P| <= 1/(eps*|M|).
    Returns the M P M residual."""
    nM = np.linalg.norm(M, 2)
    nP = np.linalg.norm(P, 2)
    res = np.linalg.norm(M @ P @ M - M, 2)
    assert res <= 4 * eps * nM
    kappa = nM * nP   # up to 1e9 at np 8
    assert np.linalg.norm(P @ M @ P - P, 2) <= 64 * np.finfo(float).eps * kappa * nP
    assert nP <= (1 + 1e-6) / (eps * nM)
    return res


def test_pinv_moore_penrose():
    for kt, np_ in [(111, 6), (301, 4), (511, 4)]:
        for which in (0, 1):
            M = O.check_matrix(kt, [1.0, 0.3], np_, which)
            eps = 0.1 ** (np_ + 1)
            check_truncated_pinv(M, O.pinv_truncated(M, eps), eps)
