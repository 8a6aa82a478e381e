"""Dense completions of a chordal partial matrix on the GPU: the drop-in for
``chompack.psdcompletion``, ``chompack.mrcompletion`` and
``chompack.edmcompletion`` (src/python/__init__.py:29,32;
pybase/psdcompletion.py, pybase/mrcompletion.py, pybase/edmcompletion.py).

The results are float64 tensors on the cspmatrix's device (the reference
returns cvxopt dense matrices): psdcompletion and edmcompletion an n x n
matrix, mrcompletion an n x r column-major matrix.  The work runs in the HIP kernels of
csrc/dense_completion.hip; there is no CPU fallback.
"""
import numpy as np

from . import _native, numeric
from .symbolic import cspmatrix

__all__ = ["psdcompletion", "mrcompletion", "edmcompletion"]


def _arrays(symb):
    return [_native.as_i64(a) for a in (symb.snptr, symb.sncolptr, symb.snrowidx, symb.blkptr)]


def _unpermute(M, symb, cols, permute_cols, perm=None):
    """M[ip, ip] (permute_cols) or M[ip, :cols]; M is column-major with
    leading dimension n (or symmetric).  ``perm`` replaces ip."""
    torch = numeric._torch()
    ip = torch.as_tensor(_native.as_i64(symb.ip if perm is None else perm), device=M.device)
    out = torch.empty_strided((symb.n, cols), (1, symb.n), dtype=torch.float64, device=M.device)
    rc = _native.lib().chp_dense_permute(M.data_ptr(), symb.n,
                                         ip.data_ptr(), symb.n, cols, int(permute_cols),
                                         out.data_ptr(), symb.n, numeric._stream(M))
    numeric._raise_rc(rc, "completion")
    torch.cuda.current_stream(M.device).synchronize()  # ip is freed on return
    return out


def _dense_sweep(fn, what, A, reordered, kwargs):
    if not (isinstance(A, cspmatrix) and A.is_factor is False):
        raise AssertionError("A must be a cspmatrix")
    torch = numeric._torch()
    tol = float(kwargs.get('tol', 1e-15))
    blk = numeric._check_blkval(A)
    symb = A.symb
    n = symb.n
    X = torch.zeros((n, n), dtype=torch.float64, device=blk.device)
    a = _arrays(symb)
    rc = fn(n, symb.Nsn, *[_native.ptr(x) for x in a], blk.data_ptr(), X.data_ptr(), n, tol,
            numeric._stream(blk))
    numeric._raise_rc(rc, what)
    if reordered:
        return X
    return _unpermute(X, symb, n, True)


def _psd_multifrontal(A, reordered):
    """The maximum-determinant completion of a positive definite completable
    A is (L L^T)^{-1} with L from the sparse completion (completion.c): the
    multifrontal completion, then trsm N and T on the identity give it in
    the original numbering.  Returns None when the completion fails (a clique
    that is not positive definite)."""
    torch = numeric._torch()
    L = A.copy()
    try:
        numeric.completion(L)
    except ArithmeticError:
        return None
    n = A.symb.n
    X = torch.eye(n, dtype=torch.float64, device=L.blkval.device).t()  # column-major
    numeric.trsm(L, X)
    numeric.trsm(L, X, trans='T')
    if not reordered:
        return X.t()
    return _unpermute(X, A.symb, n, True, perm=A.symb.p).t()


def psdcompletion(A, reordered=True, **kwargs):
    """Maximum determinant positive semidefinite matrix completion
    (pybase/psdcompletion.py:4-70): returns the dense X with P(X) = A, in the
    reordered numbering if ``reordered`` (else X[ip, ip]).

    A positive definite completable A takes the multifrontal route,
    X = (L L^T)^{-1} with L from ``completion``.  Otherwise (a clique that is
    not positive definite) the reference's sweep runs: per supernode,
    X[eta, nu] = X[eta, alpha] X[alpha, alpha]^+ X[alpha, nu] with the
    Cholesky factor of the separator block, or its eigendecomposition
    pseudo-inverse where that factorization fails (keyword ``tol``, default
    1e-15: eigenvalues <= tol * lambda_max are dropped).  Keyword
    ``sweep=True`` takes the sweep in every case."""
    if not (isinstance(A, cspmatrix) and A.is_factor is False):
        raise AssertionError("A must be a cspmatrix")
    if not kwargs.get('sweep', False):
        X = _psd_multifrontal(A, reordered)
        if X is not None:
            return X
    return _dense_sweep(_native.lib().chp_psdcompletion, "psdcompletion", A, reordered, kwargs)


def edmcompletion(A, reordered=True, **kwargs):
    """Euclidean distance matrix completion (pybase/edmcompletion.py:4-74):
    returns the dense EDM X with P(X) = A of an EDM-completable A, in the
    reordered numbering if ``reordered`` (else X[ip, ip]).  Keyword ``tol``
    as in psdcompletion (every separator takes the pseudo-inverse)."""
    return _dense_sweep(_native.lib().chp_edmcompletion, "edmcompletion", A, reordered, kwargs)


def mrcompletion(A, reordered=True):
    """Minimum rank positive semidefinite completion
    (pybase/mrcompletion.py:6-108): returns Y (n x r, r = the largest
    numerical rank of a clique) with P(Y Y^T) = A, in the reordered
    numbering if ``reordered`` (else Y[ip, :])."""
    if not (isinstance(A, cspmatrix) and A.is_factor is False):
        raise AssertionError("A must be a cspmatrix")
    torch = numeric._torch()
    blk = numeric._check_blkval(A)
    symb = A.symb
    n, maxr = symb.n, int(symb.clique_number)
    Y = torch.empty_strided((n, maxr), (1, n), dtype=torch.float64, device=blk.device)
    a = _arrays(symb)
    snpar = _native.as_i64(symb.parent())
    rank = np.zeros(1, dtype=np.int32)
    rc = _native.lib().chp_mrcompletion(n, symb.Nsn, *[_native.ptr(x) for x in a],
                                        _native.ptr(snpar), blk.data_ptr(), Y.data_ptr(), n,
                                        maxr, _native.ptr(rank), numeric._stream(blk))
    numeric._raise_rc(rc, "mrcompletion")
    r = int(rank[0])
    if reordered:
        return Y[:, :r]
    return _unpermute(Y, symb, r, False)
