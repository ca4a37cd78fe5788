"""torch.autograd binding of the FMM operator.

``fmm_apply(f, den, pos=None)`` returns ``phi = A(x) den`` for the set-up handle
``f`` (``kifmm.FMM3d``), with ``den`` a float64 tensor on the GPU (srcDOF values
per source, caller's order).  Gradients:

* with respect to the densities: the transposed pass,
  ``grad_den = A^T grad_phi`` (``kifmm_evaluate_adjoint_device``), every kernel;
* with respect to the point positions ``pos`` (an (N, 3) float64 GPU tensor
  holding the handle's points, sources = targets, Laplace single layer): row i
  and column i of A both move with x_i, and grad_s K(t, s) = grad_t K(s, t) for
  the symmetric kernel, so
  ``d<g, A den>/dx_i = g_i grad(A den)(x_i) + den_i grad(A g)(x_i)``:
  two field passes.  The first reads the upward equivalent densities that the
  forward pass checkpointed (``kifmm_evaluate_checkpoint_device``), so it skips
  the up pass.

The tree is the one built at setup: ``pos`` must hold the handle's points
(it carries the gradient, not new coordinates).  All passes are enqueued on
torch's current stream.
"""
import numpy as np
import torch


def _stream():
    return torch.cuda.current_stream().cuda_stream


class _FMMApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, den, pos, f):
        if not den.is_cuda or den.dtype != torch.float64:
            raise TypeError("densities must be a float64 GPU tensor")
        if den.numel() != f.n_src * f.src_dof:
            raise ValueError("densities must hold %d values, got %d" % (f.n_src * f.src_dof, den.numel()))
        den = den.contiguous()
        out = torch.empty(f.n_trg * f.trg_dof, dtype=torch.float64, device=den.device)
        ue = None
        if pos is not None:
            if f.kernel_type != 111 or f.n_src != f.n_trg:
                raise ValueError("position gradients need the Laplace single layer with sources = targets")
            if tuple(pos.shape) != (f.n_src, 3):
                raise ValueError("pos must be (%d, 3)" % f.n_src)
            ue = torch.empty(f.checkpoint_size(), dtype=torch.float64, device=den.device)
            f.evaluate_checkpoint_device(den, out, ue, stream=_stream())
        else:
            f.evaluate_device(den, out, stream=_stream())
        ctx.f = f
        ctx.has_pos = pos is not None
        ctx.save_for_backward(den, ue if ue is not None else torch.empty(0, device=den.device))
        return out

    @staticmethod
    def backward(ctx, grad_out):
        f = ctx.f
        den, ue = ctx.saved_tensors
        g = grad_out.to(torch.float64).contiguous()
        grad_den = grad_pos = None
        if ctx.needs_input_grad[0]:
            grad_den = torch.empty(f.n_src * f.src_dof, dtype=torch.float64, device=g.device)
            f.evaluate_adjoint_device(g, grad_den, stream=_stream())
        if ctx.has_pos and ctx.needs_input_grad[1]:
            fs = torch.empty(3 * f.n_trg, dtype=torch.float64, device=g.device)
            fg = torch.empty(3 * f.n_trg, dtype=torch.float64, device=g.device)
            f.field_checkpoint_device(den, ue, fs, stream=_stream())
            f.field_device(g, fg, stream=_stream())
            grad_pos = g[:, None] * fs.view(-1, 3) + den[:, None] * fg.view(-1, 3)
        return grad_den, grad_pos, None


def fmm_apply(f, den, pos=None):
    """phi = A den with autograd support: the transposed pass for den, two
    field passes (one from the forward pass's checkpoint) for pos."""
    return _FMMApply.apply(den, pos, f)


def position_grad(f, u, g):
    """d<g, A(x) u>/dx for sources = targets and the Laplace single layer, on
    host arrays: g_i grad(A u)(x_i) + u_i grad(A g)(x_i) (two field passes).
    u, g: float64 arrays (one value per point); returns (N, 3).  For
    L = 1/2 |A sigma - A tau|^2 use u = sigma - tau, g = A u."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    return g[:, None] * f.field(u) + u[:, None] * f.field(g)
