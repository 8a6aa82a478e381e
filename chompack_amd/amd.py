"""Approximate minimum degree ordering, the counterpart of ``cvxopt.amd.order``
(used by the reference through ``symbolic(A, p=amd.order)``, symbolic.py:540-546).

``order(A)`` reads the lower triangle of the square sparse matrix A (as
cvxopt's ``order(A, uplo='L')`` does) and returns the permutation as an int64
numpy vector: ``A[p, p]`` has a sparser Cholesky factor than A.
"""
import numpy as np
import scipy.sparse as sp

from . import _native
from .misc import symmetrize

__all__ = ["order"]


def order(A, uplo="L"):
    from .misc import as_sparse, is_sparse
    assert is_sparse(A), "A must be a square sparse matrix"
    A = as_sparse(A)
    assert A.shape[0] == A.shape[1], "A must be a square sparse matrix"
    if uplo == "U":
        A = sp.csc_matrix(A).T
    S = sp.csc_matrix(symmetrize(A))
    S.sort_indices()
    n = S.shape[0]
    cp = _native.as_i64(S.indptr)
    ri = _native.as_i64(S.indices)
    p = np.empty(n, dtype=np.int64)
    rc = _native.lib().chp_amd_order(n, _native.ptr(cp), _native.ptr(ri), _native.ptr(p))
    if rc != 0:
        raise ArithmeticError("amd ordering failed")
    return p
