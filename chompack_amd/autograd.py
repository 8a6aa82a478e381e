"""Differentiable log-determinant of a chordal positive definite matrix.

``logdet(X)`` takes a ``cspmatrix`` whose ``blkval`` (GPU float64 tensor)
holds X (lower triangle meaningful) and returns log det X as a 0-d tensor
that torch.autograd differentiates.  Forward: supernodal Cholesky of a copy
of blkval and 2 sum log diag L.  Backward: the projected inverse
Y = P(X^{-1}) of that factor (the gradient of log det restricted to the
pattern, the relation behind the barrier's gradient -Y in CHOMPACK's
src/python/pybase/hessian.py docstring), mapped to blkval coordinates: an
off-diagonal lower entry x_ij stands for both x_ij and x_ji, so
d logdet / d x_ij = 2 Y_ij, and d logdet / d x_ii = Y_ii; entries above the
diagonal inside the diagonal blocks are not part of X and get 0.
"""
import numpy as np

from .symbolic import cspmatrix
from . import numeric

__all__ = ["logdet", "LogDet"]


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


def _function():
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

    return LogDet


LogDet = None


def logdet(X):
    """log det X for a cspmatrix X (not a factor), differentiable in X.blkval."""
    global LogDet
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    if LogDet is None:
        LogDet = _function()
    return LogDet.apply(X.blkval, X.symb)
