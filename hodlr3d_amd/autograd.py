"""Differentiable products and solves with the assembled operator (caller side of
the reference's MatVecProduct / classGMRES, src/HODLR3D.cpp:21-27,
src/gmres.cpp:40-94; the reference itself has no adjoint).

With A = diag(c) + alpha * K (``HODLR3D.set_operator`` / ``set_diag``):

* ``matvec(H, b)``: y = A b; dL/db = A^T dL/dy (``h3d_matvec_transpose``).
* ``solve(H, b, c)``: x = A(c)^{-1} b; with lambda = A^{-T} dL/dx (a transposed
  refined solve), dL/db = lambda and dL/dc = -lambda * x, since dA/dc_i = e_i e_i^T.

Both run on the GPU through the C ABI; tensors must be CUDA float64.
"""
import torch


class _MatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, H, prec):
        ctx.H, ctx.prec = H, prec
        return H.matvec(b, prec=prec)

    @staticmethod
    def backward(ctx, gy):
        gb = ctx.H.matvec(gy.contiguous(), prec=ctx.prec, transpose=True)
        return gb, None, None


def matvec(H, b, prec="fp64"):
    """y = A b, differentiable in b."""
    return _MatVec.apply(b, H, prec)


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, c, H, opts):
        if c is not None:
            H.set_diag(c)
        x, info = H.solve(b, **opts)
        ctx.H, ctx.opts = H, opts
        ctx.save_for_backward(x, c if c is not None else torch.empty(0))
        ctx.has_c = c is not None
        return x

    @staticmethod
    def backward(ctx, gx):
        x, c = ctx.saved_tensors
        H = ctx.H
        if ctx.has_c:
            H.set_diag(c)  # the operator may have been changed since the forward pass
        lam, _ = H.solve(gx.contiguous(), transpose=True, **ctx.opts)
        gc = -lam * x if ctx.has_c and ctx.needs_input_grad[1] else None
        return lam, gc, None, None


def solve(H, b, c=None, tol=1e-12, restart=30, max_inner=600, inner_tol=1e-6, inner_prec="fp64",
          max_outer=20):
    """x = A^{-1} b with A = diag(c) + alpha K (c: per-point diagonal in the
    original order, or None for the operator as set); differentiable in b and c."""
    opts = dict(tol=tol, restart=restart, max_inner=max_inner, inner_tol=inner_tol,
                inner_prec=inner_prec, max_outer=max_outer)
    return _Solve.apply(b, c, H, opts)
