"""Sparse-pattern helpers and the frontal extend-add / extract primitives.

Host-side mirror of chompack/misc.py (src/python/misc.py:4-204): ``tril``,
``triu``, ``symmetrize``, ``perm``, ``eye``, ``lmerge``, and the device
versions of ``frontal_add_update`` / ``frontal_get_update`` (cbase.c:11-59).
Sparse matrices are ``scipy.sparse`` matrices; a cvxopt-style ``spmatrix``
(an object with ``size`` and ``CCS``, or ``size`` and ``I``/``J``/``V``) is
accepted wherever a sparse matrix is and converted by ``as_sparse``.
Explicit zeros are kept as structural entries, as cvxopt keeps them.
"""
import numpy as np
import scipy.sparse as sp

from . import _native

__all__ = ["tril", "triu", "symmetrize", "perm", "eye", "lmerge",
           "frontal_add_update", "frontal_get_update", "coo_triplets", "as_sparse",
           "is_sparse"]


def _vec(v):
    """1-D numpy copy of a sequence or buffer (a cvxopt ``matrix`` column)."""
    return np.asarray(v).reshape(-1, order="F")


def is_sparse(A):
    """True for a scipy sparse matrix or a cvxopt-style ``spmatrix``."""
    return sp.issparse(A) or (hasattr(A, "size") and (
        hasattr(A, "CCS") or (hasattr(A, "I") and hasattr(A, "J") and hasattr(A, "V"))))


def as_sparse(A):
    """scipy CSC matrix of a sparse argument.

    scipy matrices pass through.  A cvxopt-style ``spmatrix`` is read through
    the attributes cvxopt exposes: ``size`` (rows, cols) and ``CCS``
    (colptr, rowind, values), or ``I``, ``J``, ``V`` triplets.  Values must
    be real."""
    if sp.issparse(A):
        return A
    if not is_sparse(A):
        raise TypeError("expected a sparse matrix")
    m, n = (int(x) for x in A.size)
    if hasattr(A, "CCS"):
        cp, ri, v = A.CCS
        cp, ri, v = _vec(cp).astype(np.int64), _vec(ri).astype(np.int64), _vec(v)
        if np.iscomplexobj(v):
            raise TypeError("complex sparse matrices are not supported")
        return sp.csc_matrix((v.astype(np.float64), ri, cp), shape=(m, n))
    I, J, V = _vec(A.I).astype(np.int64), _vec(A.J).astype(np.int64), _vec(A.V)
    if np.iscomplexobj(V):
        raise TypeError("complex sparse matrices are not supported")
    return sp.csc_matrix((V.astype(np.float64), (I, J)), shape=(m, n))


def coo_triplets(A):
    """(I, J, V) of a sparse matrix with duplicates summed, column-major sorted.

    Explicit zeros are kept (they are part of the sparsity pattern)."""
    A = sp.coo_matrix(as_sparse(A))
    n0, n1 = A.shape
    I = np.asarray(A.row, dtype=np.int64)
    J = np.asarray(A.col, dtype=np.int64)
    V = np.asarray(A.data, dtype=np.float64)
    key = J * n0 + I
    order = np.argsort(key, kind="stable")
    key = key[order]
    V = V[order]
    if key.size:
        uniq, start = np.unique(key, return_index=True)
        V = np.add.reduceat(V, start) if V.size else V
        key = uniq
    return key % n0, key // n0, V


def _from_triplets(I, J, V, shape):
    return sp.csc_matrix((V, (I, J)), shape=shape)


def tril(A):
    """Lower triangular part of a sparse or dense matrix (misc.py:4-16)."""
    if is_sparse(A):
        A = as_sparse(A)
        I, J, V = coo_triplets(A)
        m = I >= J
        return _from_triplets(I[m], J[m], V[m], A.shape)
    return np.tril(np.asarray(A))


def triu(A):
    """Upper triangular part of a sparse or dense matrix (misc.py:18-30)."""
    if is_sparse(A):
        A = as_sparse(A)
        I, J, V = coo_triplets(A)
        m = I <= J
        return _from_triplets(I[m], J[m], V[m], A.shape)
    return np.triu(np.asarray(A))


def symmetrize(A):
    """Symmetric matrix from the lower triangle of a square sparse matrix
    (misc.py:32-40): tril(A) + strictly-lower(A)^T."""
    assert is_sparse(A), "argument must be a sparse matrix"
    A = as_sparse(A)
    assert A.shape[0] == A.shape[1], "argument must me a square matrix"
    I, J, V = coo_triplets(A)
    lo = I >= J
    st = I > J
    I2 = np.concatenate([I[lo], J[st]])
    J2 = np.concatenate([J[lo], I[st]])
    V2 = np.concatenate([V[lo], V[st]])
    return _from_triplets(I2, J2, V2, A.shape)


def perm(A, p):
    """Symmetric permutation A[p, p] of a sparse matrix (misc.py:42-53)."""
    assert is_sparse(A), "argument must be a sparse matrix"
    A = as_sparse(A)
    assert A.shape[0] == A.shape[1], "A must be a square matrix"
    p = np.asarray(p, dtype=np.int64).ravel()
    assert A.shape[0] == len(p), "length of p must be equal to the order of A"
    ip = np.empty_like(p)
    ip[p] = np.arange(len(p), dtype=np.int64)
    I, J, V = coo_triplets(A)
    return _from_triplets(ip[I], ip[J], V, A.shape)


def eye(n):
    """Dense identity matrix of order n (misc.py:55-61)."""
    return np.eye(n)


def lmerge(left, right, offsetl=0, offsetr=0, nl=None, nr=None):
    """Merge the ascending segments left[offsetl:offsetl+nl] and
    right[offsetr:offsetr+nr] and write the result to left[offsetl:...]
    (cbase.c:64-101, misc.py:178-204).  Equal heads are taken once, so a value
    repeated inside one segment stays repeated.  ``left`` is any mutable
    integer sequence (list, numpy array, cvxopt ``matrix``); it must have room
    for the merged segment.  Returns the merged length."""
    if nl is None:
        nl = len(left)
    if nr is None:
        nr = len(right)
    nl, nr, offsetl, offsetr = int(nl), int(nr), int(offsetl), int(offsetr)
    if min(nl, nr, offsetl, offsetr) < 0:
        raise ValueError("offsets and lengths must be non-negative")
    a = np.ascontiguousarray(_vec(left)[offsetl:offsetl + nl], dtype=np.int64)
    b = np.ascontiguousarray(_vec(right)[offsetr:offsetr + nr], dtype=np.int64)
    if a.size != nl or b.size != nr:
        raise ValueError("segment extends past the end of its list")
    out = np.empty(nl + nr, dtype=np.int64)
    k = int(_native.lib().chp_lmerge(_native.ptr(a), _native.ptr(b), nl, nr, _native.ptr(out)))
    if offsetl + k > len(left):
        raise ValueError("left has no room for the merged list (%d entries from %d)"
                         % (k, offsetl))
    if isinstance(left, list):
        left[offsetl:offsetl + k] = [int(x) for x in out[:k]]
    elif isinstance(left, np.ndarray):
        left[offsetl:offsetl + k] = out[:k]
    else:
        for i in range(k):
            left[offsetl + i] = int(out[i])
    return k


def _strides_of(T):
    import torch
    assert isinstance(T, torch.Tensor) and T.dim() == 2
    assert T.dtype == torch.float64 and T.is_cuda, "expected a float64 CUDA tensor"
    return T.stride(0), T.stride(1)


def _rel_rows(F, relidx, relptr, i):
    """Host int64 copy of r = relidx[relptr[i]:relptr[i+1]], checked against
    F: every index a row and column of F, no index twice (one GPU thread
    updates each target entry, so a repeated index would race)."""
    i = int(i)
    rp = _vec(relptr)
    if not 0 <= i < rp.size - 1:
        raise IndexError("supernode index out of range")
    r = _vec(relidx)[int(rp[i]):int(rp[i + 1])].astype(np.int64)
    if r.size and (r.min() < 0 or r.max() >= min(F.shape[0], F.shape[1])):
        raise IndexError("relative index outside the frontal matrix")
    if np.unique(r).size != r.size:
        raise ValueError("relative indices must be distinct")
    return r


def frontal_add_update(F, U, relidx, relptr, i, alpha=1.0):
    """F[r, r] += alpha * lower(U) with r = relidx[relptr[i]:relptr[i+1]]
    (cbase.c:11-34): only U's lower triangle is added, as in the reference's
    C version.  F and U are float64 CUDA tensors of any strides, F updated in
    place on the current stream.  U must be N x N with N = len(r) (the
    reference returns an error there too)."""
    import torch
    lib = _native.lib()
    fr, fc = _strides_of(F)
    ur, uc = _strides_of(U)
    r = _rel_rows(F, relidx, relptr, i)
    N = len(r)
    if U.shape[0] != N or U.shape[1] != N:
        raise ValueError("U must be %d x %d" % (N, N))
    if U.device != F.device:
        raise ValueError("F and U must be on the same device")
    d_r = torch.as_tensor(r, device=F.device)
    stream = torch.cuda.current_stream(F.device).cuda_stream
    st = (ctypes_i64 * 4)(fr, fc, ur, uc)
    rc = lib.chp_frontal_add_update(F.data_ptr(), N, d_r.data_ptr(), U.data_ptr(), st,
                                    float(alpha), stream)
    _check(rc)
    return None


def frontal_get_update(F, relidx, relptr, i):
    """Return U = lower(F[r, r]) as a new N x N float64 CUDA tensor
    (cbase.c:39-59); the strict upper part of U is zero.  F may have any
    strides."""
    import torch
    lib = _native.lib()
    fr, fc = _strides_of(F)
    r = _rel_rows(F, relidx, relptr, i)
    N = len(r)
    U = torch.zeros((N, N), dtype=torch.float64, device=F.device)
    d_r = torch.as_tensor(r, device=F.device)
    stream = torch.cuda.current_stream(F.device).cuda_stream
    st = (ctypes_i64 * 4)(fr, fc, U.stride(0), U.stride(1))
    rc = lib.chp_frontal_get_update(F.data_ptr(), N, d_r.data_ptr(),
                                    U.data_ptr(), st, stream)
    _check(rc)
    return U


import ctypes as _ct  # noqa: E402
ctypes_i64 = _ct.c_int64


def _check(rc):
    if rc != 0:
        msg = _native.lib().chp_last_error()
        raise RuntimeError("chompack_amd: %s (code %d)" % (msg.decode() if msg else "error", rc))
