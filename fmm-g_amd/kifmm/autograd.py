"""torch.autograd binding of the FMM operator.

``fmm_apply(f, den)`` returns ``phi = A den`` for the set-up handle ``f``
(``kifmm.FMM3d``), with ``den`` a float64 tensor on the GPU (srcDOF values per
source, caller's order).  Its backward is the transposed pass,
``grad_den = A^T grad_phi`` (``kifmm_evaluate_adjoint_device``), so losses built
from FMM potentials differentiate with respect to the densities.  Both passes
are enqueued on torch's current stream.

Position gradients are not part of the autograd graph: for sources equal to
targets and the Laplace single layer, dL/dx_i = g_i grad(A sigma)_i -
sigma_i grad(A g)_i with g = dL/dphi, i.e. two ``FMM3d.field`` passes.
"""
import torch


def _stream():
    return torch.cuda.current_stream().cuda_stream


class _FMMApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, den, f):
        if not den.is_cuda or den.dtype != torch.float64:
            raise TypeError("densities must be a float64 GPU tensor")
        if den.numel() != f.n_src * f.src_dof:
            raise ValueError("densities must hold %d values, got %d" % (f.n_src * f.src_dof, den.numel()))
        den = den.contiguous()
        out = torch.empty(f.n_trg * f.trg_dof, dtype=torch.float64, device=den.device)
        f.evaluate_device(den, out, stream=_stream())
        ctx.f = f
        return out

    @staticmethod
    def backward(ctx, grad_out):
        f = ctx.f
        g = grad_out.to(torch.float64).contiguous()
        out = torch.empty(f.n_src * f.src_dof, dtype=torch.float64, device=g.device)
        f.evaluate_adjoint_device(g, out, stream=_stream())
        return out, None


def fmm_apply(f, den):
    """phi = A den with autograd support through the transposed pass."""
    return _FMMApply.apply(den, f)


def position_grad(f, u, g):
    """d<g, A(x) u>/dx for sources = targets and the Laplace single layer.

    Every point x_i is both a target (row i) and a source (column i), and
    grad_s K(t, s) = grad_t K(s, t) for the symmetric kernel, so
    d<g, A u>/dx_i = g_i grad(A u)(x_i) + u_i grad(A g)(x_i):
    two field passes.  u, g: float64 arrays (one value per point); returns
    (N, 3).  For L = 1/2 |A sigma - A tau|^2 use u = sigma - tau, g = A u.
    """
    import numpy as np
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    return g[:, None] * f.field(u) + u[:, None] * f.field(g)
