"""Product-form Cholesky factorization on the GPU (pfcholesky.py:113-211,
pfcholesky.c).

    X + V diag(a) V^T = L_0 L_1 ... L_m L_m^T ... L_1^T L_0^T

with X = L_0 L_0^T a chordal Cholesky factorization and L_i the product-form
factors of I + Vt diag(a) Vt^T, Vt = L_0^{-1} V.  The factors live on the
GPU as column-major n x m float64 tensors; ``pfchol``, ``pftrsm`` and
``pftrmm`` mirror the reference's cbase functions of the same names
(cbase.c:859-942) and run the scan kernels of csrc/pf.hip.
"""
import numpy as np
import scipy.sparse as sp

from . import numeric, _native
from .misc import as_sparse, is_sparse
from .symbolic import cspmatrix, symbolic

__all__ = ["pfcholesky", "pfchol", "pftrsm", "pftrmm"]


def _colmajor(T, name):
    """(work, rows, cols, ld, copied) of a float64 GPU vector or matrix:
    ``work`` is T, or a column-major copy of it when T's strides are not
    column-major (written back by the caller)."""
    import torch
    if not (isinstance(T, torch.Tensor) and T.is_cuda and T.dtype == torch.float64):
        raise TypeError("%s must be a float64 tensor in GPU memory" % name)
    if T.dim() not in (1, 2):
        raise TypeError("%s must be a vector or a matrix" % name)
    work, ld, _, copied = numeric.colmajor_work(T)
    return work, T.shape[0], (1 if T.dim() == 1 else T.shape[1]), ld, copied


def _need(T, name, n, m):
    if T[1] != n or T[2] < m:
        raise ValueError("%s must have %d rows and at least %d columns" % (name, n, m))


def _write_back(pairs):
    for T, w in pairs:
        if w[4]:
            T.copy_(w[0])


def pfchol(a, V, L, B):
    """Product-form factorization of I + V diag(a) V^T in place: on exit
    L and B hold the factors and V's columns 1.. are transformed as in
    dpftrf (pfcholesky.c:52-87).  L and B need V's row count and at least
    its column count.  Raises ArithmeticError("factorization failed") where
    the reference does."""
    wv, wl, wb = _colmajor(V, "V"), _colmajor(L, "L"), _colmajor(B, "B")
    n, m = wv[1], wv[2]
    _need(wl, "L", n, m)
    _need(wb, "B", n, m)
    ah = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()[:m])
    if ah.size < m:
        raise ValueError("a must have at least as many entries as V has columns")
    rc = _native.lib().chp_pf_factor(n, m, _native.ptr(ah), wv[0].data_ptr(), wv[3],
                                     wl[0].data_ptr(), wl[3], wb[0].data_ptr(), wb[3],
                                     numeric._stream(V))
    numeric._raise_rc(rc, "pfchol")
    _write_back([(V, wv), (L, wl), (B, wb)])
    if rc == 1:
        raise ArithmeticError("factorization failed")


def _pf_apply(fn, what, V, L, B, U, trans):
    if trans not in ('N', 'T'):
        raise ValueError("trans must be 'N' or 'T'")
    wv, wl, wb, wu = _colmajor(V, "V"), _colmajor(L, "L"), _colmajor(B, "B"), _colmajor(U, "U")
    n, m = wv[1], wv[2]
    _need(wl, "L", n, m)
    _need(wb, "B", n, m)
    if wu[1] != n:
        raise ValueError("U must have as many rows as V")
    rc = fn(n, m, wv[0].data_ptr(), wv[3], wl[0].data_ptr(), wl[3], wb[0].data_ptr(), wb[3],
            wu[0].data_ptr(), wu[3], wu[2], ord(trans), numeric._stream(U))
    numeric._raise_rc(rc, what)
    _write_back([(U, wu)])


def pftrsm(V, L, B, U, trans='N'):
    """U := (L_1 ... L_m)^{-1} U ('N') or its transpose ('T'), dpfsv
    (pfcholesky.c:89-112) for every column of U."""
    _pf_apply(_native.lib().chp_pf_solve, "pftrsm", V, L, B, U, trans)


def pftrmm(V, L, B, U, trans='N'):
    """U := L_1 ... L_m U ('N') or its transpose ('T'), dpfmv
    (pfcholesky.c:114-136) for every column of U."""
    _pf_apply(_native.lib().chp_pf_product, "pftrmm", V, L, B, U, trans)


class pfcholesky(object):
    """Supernodal multifrontal product-form Cholesky factorization
    X + V diag(a) V^T = L_m ... L_1 L_0 L_0^T L_1^T ... L_m^T
    (pfcholesky.py:113-211).

    :param X: cspmatrix (not a factor: it is copied and factored; a factor:
              used as L_0) or a scipy sparse matrix (lower triangle; a
              symbolic factorization with ordering p is made)
    :param V: n x m float64 matrix (numpy or GPU tensor)
    :param a: m weights (default: ones)
    :param p: ordering for a sparse X (default: natural ordering)
    """

    def __init__(self, X, V, a=None, p=None, device="cuda:0"):
        import torch
        if isinstance(X, cspmatrix):
            n = X.symb.n
            device = X.blkval.device
        elif is_sparse(X):
            X = as_sparse(X)
            n = X.shape[0]
            if X.shape[0] != X.shape[1]:
                raise ValueError("X must be a square matrix")
        else:
            raise TypeError("X must be a cspmatrix or a sparse matrix")
        if not isinstance(V, torch.Tensor):
            V = np.asarray(V, dtype=np.float64)   # numpy, or a buffer such as a cvxopt matrix
        Vt = torch.as_tensor(V, dtype=torch.float64).to(device)
        if Vt.dim() == 1:
            Vt = Vt.reshape(-1, 1)
        if Vt.shape[0] != n:
            raise ValueError("V must have the same number of rows as X")
        self._n = n
        m = Vt.shape[1]
        self._V = torch.empty_strided((n, m), (1, n), dtype=torch.float64, device=device)
        self._V.copy_(Vt)
        if isinstance(X, cspmatrix) and X.is_factor is False:
            self._L0 = X.copy()
            numeric.cholesky(self._L0)
        elif isinstance(X, cspmatrix):
            self._L0 = X
        else:
            symb = symbolic(X, p=p)
            self._L0 = cspmatrix(symb, device=device) + X
            numeric.cholesky(self._L0)
        numeric.trsm(self._L0, self._V)
        if a is None:
            a = np.ones(m)
        self._L = torch.zeros_like(self._V)
        self._B = torch.zeros_like(self._V)
        pfchol(a, self._V, self._L, self._B)

    def __repr__(self):
        return "<%ix%i product-form Cholesky factor, r=%i, tc='d'>" % (
            self._n, self._n, self._V.shape[1])

    def trsm(self, B, trans='N'):
        """B := L^{-1} B ('N') or L^{-T} B ('T') with the product-form factor."""
        if not isinstance(trans, str):
            raise TypeError("trans must be 'N' or 'T'")
        if trans == 'N':
            numeric.trsm(self._L0, B)
            pftrsm(self._V, self._L, self._B, B, trans='N')
        elif trans == 'T':
            pftrsm(self._V, self._L, self._B, B, trans='T')
            numeric.trsm(self._L0, B, trans='T')
        else:
            raise ValueError("trans must be 'N' or 'T'")

    def trmm(self, B, trans='N'):
        """B := L B ('N') or L^T B ('T') with the product-form factor."""
        if not isinstance(trans, str):
            raise TypeError("trans must be 'N' or 'T'")
        if trans == 'N':
            pftrmm(self._V, self._L, self._B, B, trans='N')
            numeric.trmm(self._L0, B)
        elif trans == 'T':
            numeric.trmm(self._L0, B, trans='T')
            pftrmm(self._V, self._L, self._B, B, trans='T')
        else:
            raise ValueError("trans must be 'N' or 'T'")
