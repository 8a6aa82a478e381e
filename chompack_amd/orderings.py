"""Orderings of chordal sparsity patterns: maximum cardinality search
(mcs.py:4-57) and the maximal chordal subgraph (maxchord.py:5-83), on the
host in C++ (csrc/symbolic.cpp); ``peo`` (symbolic.py:413-449) is in
``symbolic``.  Integer results are the reference's for the same pattern."""
import numpy as np
import scipy.sparse as sp

from . import _native
from .misc import symmetrize, as_sparse, is_sparse

__all__ = ["maxcardsearch", "maxchord"]


def _pattern(A, ve):
    if not is_sparse(A):
        raise TypeError("A must be a sparse matrix")
    A = as_sparse(A)
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("A must be a square matrix")
    if ve is None:
        ve = n - 1
    elif not (isinstance(ve, (int, np.integer)) and 0 <= ve < n):
        raise ValueError("ve must be an integer between 0 and A.size[0]-1")
    As = sp.csc_matrix(symmetrize(A))
    As.sort_indices()
    return n, int(ve), As


def maxcardsearch(A, ve=None):
    """Maximum cardinality search ordering of a chordal matrix (only the lower
    triangle of A is read); ``ve`` is the last vertex eliminated (default
    n - 1).  The result is a perfect elimination ordering of a chordal A."""
    n, ve, As = _pattern(A, ve)
    cp = np.ascontiguousarray(As.indptr, dtype=np.int64)
    ri = np.ascontiguousarray(As.indices, dtype=np.int64)
    p = np.empty(n, dtype=np.int64)
    _native.lib().chp_maxcardsearch(n, _native.ptr(cp), _native.ptr(ri), ve, _native.ptr(p))
    return p


def maxchord(A, ve=None):
    """Maximal chordal subgraph: returns (Ac, p) with Ac the lower-triangular
    projection of A on a maximal chordal subgraph of its sparsity graph and p
    a perfect elimination ordering of Ac (Dearing, Shier and Warner, 1988).
    If A is chordal, Ac has A's pattern and values."""
    n, ve, As = _pattern(A, ve)
    cp = np.ascontiguousarray(As.indptr, dtype=np.int64)
    ri = np.ascontiguousarray(As.indices, dtype=np.int64)
    nz = int(cp[-1])
    p = np.empty(n, dtype=np.int64)
    rows = np.empty(max(nz, 1), dtype=np.int64)
    cols = np.empty(max(nz, 1), dtype=np.int64)
    src = np.empty(max(nz, 1), dtype=np.int64)
    cnt = _native.lib().chp_maxchord(n, _native.ptr(cp), _native.ptr(ri), ve, _native.ptr(p),
                                     _native.ptr(rows), _native.ptr(cols), _native.ptr(src))
    vals = np.asarray(As.data, dtype=np.float64)[src[:cnt]]
    Ac = sp.csc_matrix((vals, (rows[:cnt], cols[:cnt])), shape=(n, n))
    return Ac, p
