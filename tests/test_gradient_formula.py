"""Position gradient of an FMM loss (the gradient workload: L = 1/2 |A(x) sigma -
A(x) tau|^2 with sources = targets, Laplace single layer).  On the CPU, with
direct sums standing in for A: the closed form dL/dx_i = g_i grad(A u)_i +
u_i grad(A g)_i (u = sigma - tau, g = A u), used by kifmm.autograd.position_grad,
against central differences with h = 1e-6 on 64 random coordinates.  Self pairs
stay excluded under the perturbation because the moved point is both the source
and the target of the pair."""
import numpy as np

from kifmm import workloads as W


def apply(x, u):
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    k = np.where(r < 1e-8, 0.0, 1.0 / np.where(r < 1e-8, 1.0, r)) / (4 * np.pi)
    return k @ u


def field(x, u):
    d = x[:, None, :] - x[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    inv3 = np.where(r2 < 1e-16, 0.0, np.where(r2 < 1e-16, 1.0, r2) ** -1.5)
    return -np.einsum("ij,j,ijk->ik", inv3, u, d) / (4 * np.pi)


def loss(x, sig, tau):
    e = apply(x, sig) - apply(x, tau)
    return 0.5 * np.sum(e * e)


def test_position_gradient_closed_form():
    n = 400
    x = W.uniform(n, W.seed(3, 0))
    sig = W.densities(n, 1, W.seed(3, 1))
    tau = W.densities(n, 1, W.seed(3, 2))
    u = sig - tau
    g = apply(x, u)
    grad = g[:, None] * field(x, u) + u[:, None] * field(x, g)
    rng = np.random.default_rng(W.seed(3, 4))
    h = 1e-6
    worst = 0.0
    for _ in range(64):
        i, k = rng.integers(n), rng.integers(3)
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd = (loss(xp, sig, tau) - loss(xm, sig, tau)) / (2 * h)
        worst = max(worst, abs(fd - grad[i, k]) / (abs(grad[i, k]) + 1e-3 * np.abs(grad).max()))
    assert worst < 1e-5, worst


def test_density_gradient_is_adjoint():
    """dL/dsigma = A^T (A u) with A symmetric here: FD on 16 densities."""
    n = 300
    x = W.uniform(n, W.seed(3, 0) + 7)
    sig = W.densities(n, 1, 5)
    tau = W.densities(n, 1, 6)
    g = apply(x, sig - tau)
    grad = apply(x, g)  # A^T = A for the symmetric single layer with sources = targets
    rng = np.random.default_rng(8)
    h = 1e-6
    for _ in range(16):
        i = rng.integers(n)
        sp, sm = sig.copy(), sig.copy()
        sp[i] += h
        sm[i] -= h
        fd = (loss(x, sp, tau) - loss(x, sm, tau)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * (abs(grad[i]) + np.abs(grad).max() * 1e-3)
