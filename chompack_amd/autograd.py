"""Differentiable chordal factorizations (torch.autograd Functions whose
backward passes are hand-written sweeps on the GPU).

All functions take a ``cspmatrix`` X whose ``blkval`` (GPU float64 tensor)
holds X's lower-stored entries, and differentiate with respect to that
tensor: an off-diagonal lower entry x_ij stands for both x_ij and x_ji, so
its gradient is twice the (i, j) entry of the symmetric gradient; entries
above the diagonal inside the diagonal blocks are not part of X and get 0.

* ``logdet(X)``: log det X.  Forward: Cholesky and 2 sum log diag L.
  Backward: the projected inverse Y = P(X^{-1}) of that factor (the
  gradient of log det restricted to the pattern; the barrier's gradient -Y
  in CHOMPACK's pybase/hessian.py docstring).
* ``cholesky(X)``: the factor's blkval (stored convention, cholesky.c:71-72).
  Backward: ``numeric.cholesky_adjoint``, the reverse-mode sweep of the
  multifrontal factorization.  With ``recompute=True`` the factor is not
  kept for the backward: the backward factors X again (fronts recomputed
  instead of stored, for callers short of device memory).
* ``completion(X)``: the factor L of the maximum-determinant completion
  (completion.c), P((L L^T)^{-1}) = X.  With Lam = L L^T (sparse on the
  pattern), X = P(Lam^{-1}), so dLam = -H^{-1}(dX) with H the Hessian map
  of the log-det barrier at Lam (hessian.c): the backward is the Cholesky
  adjoint sweep at Lam followed by -hessian(L, X, ., adj=None, inv=True).
* ``maxdet(X)``: the maximum log det of a completion of X,
  -2 sum log diag L.  Its gradient is Lam = L L^T on the pattern (the
  Lagrange multiplier of P(Z) = X), the llt sweep (llt.c).
"""
import numpy as np

from .symbolic import cspmatrix
from . import numeric

__all__ = ["logdet", "cholesky", "completion", "maxdet"]


def _grad_weights(symb, device):
    """2 on strictly-lower entries, 1 on the diagonal, 0 above it (per block)."""
    import torch
    key = "_logdet_w_%s" % device
    w = getattr(symb, key, None)
    if w is None:
        h = np.zeros(int(symb.blkptr[-1]))
        for k in range(symb.Nsn):
            nn = int(symb.snptr[k + 1] - symb.snptr[k])
            nj = nn + int(symb.relptr[k + 1] - symb.relptr[k])
            blk = h[symb.blkptr[k]:symb.blkptr[k + 1]].reshape((nn, nj)).T
            ii, jj = np.tril_indices(nj, -1, nn)
            blk[ii, jj] = 2.0
            blk[np.arange(nn), np.arange(nn)] = 1.0
        w = torch.as_tensor(h, device=device)
        setattr(symb, key, w)
    return w


_F = {}


def _functions():
    if _F:
        return _F
    import torch

    class LogDet(torch.autograd.Function):
        @staticmethod
        def forward(ctx, blkval, symb):
            L = cspmatrix(symb, blkval=blkval.detach().clone())
            numeric.cholesky(L)
            val = numeric.logdet(L)
            ctx.symb = symb
            ctx.save_for_backward(L.blkval)
            return blkval.new_tensor(val)

        @staticmethod
        def backward(ctx, g):
            (lb,) = ctx.saved_tensors
            Y = cspmatrix(ctx.symb, blkval=lb.clone(), factor=True)
            numeric.projected_inverse(Y)
            return g * Y.blkval * _grad_weights(ctx.symb, lb.device), None

    class Cholesky(torch.autograd.Function):
        @staticmethod
        def forward(ctx, blkval, symb, recompute):
            L = cspmatrix(symb, blkval=blkval.detach().clone())
            numeric.cholesky(L)
            ctx.symb, ctx.recompute = symb, recompute
            ctx.save_for_backward(blkval if recompute else L.blkval)
            return L.blkval

        @staticmethod
        def backward(ctx, g):
            (t,) = ctx.saved_tensors
            symb = ctx.symb
            L = cspmatrix(symb, blkval=t.detach().clone())
            if ctx.recompute:
                numeric.cholesky(L)
            else:
                L.is_factor = True
            G = cspmatrix(symb, blkval=g.detach().contiguous())
            Fb = numeric.cholesky_adjoint(L, G)
            return Fb.blkval * _grad_weights(symb, t.device), None, None

    class Completion(torch.autograd.Function):
        @staticmethod
        def forward(ctx, blkval, symb, factored):
            L = cspmatrix(symb, blkval=blkval.detach().clone())
            numeric.completion(L, factored_updates=factored)
            ctx.symb, ctx.factored = symb, factored
            ctx.save_for_backward(blkval, L.blkval)
            return L.blkval

        @staticmethod
        def backward(ctx, g):
            xb, lb = ctx.saved_tensors
            symb = ctx.symb
            L = cspmatrix(symb, blkval=lb.detach(), factor=True)
            Fb = numeric.cholesky_adjoint(L, cspmatrix(symb, blkval=g.detach().contiguous()))
            Y = cspmatrix(symb, blkval=xb.detach())
            numeric.hessian(L, Y, [Fb], adj=None, inv=True, factored_updates=ctx.factored)
            return -Fb.blkval * _grad_weights(symb, xb.device), None, None

    class MaxDet(torch.autograd.Function):
        @staticmethod
        def forward(ctx, blkval, symb, factored):
            L = cspmatrix(symb, blkval=blkval.detach().clone())
            numeric.completion(L, factored_updates=factored)
            val = -numeric.logdet(L)
            ctx.symb = symb
            ctx.save_for_backward(L.blkval)
            return blkval.new_tensor(val)

        @staticmethod
        def backward(ctx, g):
            (lb,) = ctx.saved_tensors
            Lam = cspmatrix(ctx.symb, blkval=lb.clone(), factor=True)
            numeric.llt(Lam)
            return g * Lam.blkval * _grad_weights(ctx.symb, lb.device), None, None

    _F.update(LogDet=LogDet, Cholesky=Cholesky, Completion=Completion, MaxDet=MaxDet)
    return _F


def _check(X):
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")


def logdet(X):
    """log det X for a cspmatrix X (not a factor), differentiable in X.blkval."""
    _check(X)
    return _functions()["LogDet"].apply(X.blkval, X.symb)


def cholesky(X, recompute=False):
    """Stored-convention Cholesky factor of X as a blkval tensor,
    differentiable in X.blkval (backward: the Cholesky adjoint sweep).
    ``recompute=True`` keeps X instead of the factor for the backward, which
    then factors X again."""
    _check(X)
    return _functions()["Cholesky"].apply(X.blkval, X.symb, bool(recompute))


def completion(X, factored_updates=True):
    """Factor L of the maximum-determinant completion of X as a blkval
    tensor, differentiable in X.blkval (backward: Cholesky adjoint sweep,
    then the inverse Hessian map at L L^T)."""
    _check(X)
    return _functions()["Completion"].apply(X.blkval, X.symb, bool(factored_updates))


def maxdet(X, factored_updates=True):
    """max log det Z over the completions Z of X (P(Z) = X), differentiable
    in X.blkval (backward: L L^T by the llt sweep)."""
    _check(X)
    return _functions()["MaxDet"].apply(X.blkval, X.symb, bool(factored_updates))
