"""Sparse-pattern helpers and the frontal extend-add / extract primitives.

Host-side mirror of chompack/misc.py (src/python/misc.py:4-204): ``tril``,
``triu``, ``symmetrize``, ``perm``, ``eye``, ``lmerge``, and the device
versions of ``frontal_add_update`` / ``frontal_get_update`` (cbase.c:11-59).
Sparse matrices are ``scipy.sparse`` matrices (cvxopt's ``spmatrix`` has no
counterpart here); explicit zeros are kept as structural entries, as cvxopt
keeps them.
"""
import numpy as np
import scipy.sparse as sp

from . import _native

__all__ = ["tril", "triu", "symmetrize", "perm", "eye", "lmerge",
           "frontal_add_update", "frontal_get_update", "coo_triplets"]


def coo_triplets(A):
    """(I, J, V) of a sparse matrix with duplicates summed, column-major sorted.

    Explicit zeros are kept (they are part of the sparsity pattern)."""
    A = sp.coo_matrix(A)
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
    if sp.issparse(A):
        I, J, V = coo_triplets(A)
        m = I >= J
        return _from_triplets(I[m], J[m], V[m], A.shape)
    return np.tril(np.asarray(A))


def triu(A):
    """Upper triangular part of a sparse or dense matrix (misc.py:18-30)."""
    if sp.issparse(A):
        I, J, V = coo_triplets(A)
        m = I <= J
        return _from_triplets(I[m], J[m], V[m], A.shape)
    return np.triu(np.asarray(A))


def symmetrize(A):
    """Symmetric matrix from the lower triangle of a square sparse matrix
    (misc.py:32-40): tril(A) + strictly-lower(A)^T."""
    assert sp.issparse(A), "argument must be a sparse matrix"
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
    assert sp.issparse(A), "argument must be a sparse matrix"
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
    """Sorted union of two sorted segments written back into ``left``
    (cbase.c:64-101, misc.py:178-204).  Returns the merged length."""
    if nl is None:
        nl = len(left)
    if nr is None:
        nr = len(right)
    a = np.asarray(left[offsetl:offsetl + nl])
    b = np.asarray(right[offsetr:offsetr + nr])
    m = np.union1d(a, b)
    left[offsetl:offsetl + len(m)] = m
    return int(len(m))


def _strides_of(T):
    import torch
    assert isinstance(T, torch.Tensor) and T.dim() == 2
    assert T.dtype == torch.float64 and T.is_cuda, "expected a float64 CUDA tensor"
    return T.stride(0), T.stride(1)


def frontal_add_update(F, U, relidx, relptr, i, alpha=1.0):
    """F[r, r] += alpha * lower(U) with r = relidx[relptr[i]:relptr[i+1]]
    (cbase.c:11-34).  F and U are float64 CUDA tensors, updated in place on
    the current stream."""
    import torch
    lib = _native.lib()
    r = np.asarray(relidx[int(relptr[i]):int(relptr[i + 1])], dtype=np.int64)
    N = len(r)
    if U.shape[0] != N:
        return None
    fr, fc = _strides_of(F)
    ur, uc = _strides_of(U)
    d_r = torch.as_tensor(r, device=F.device)
    stream = torch.cuda.current_stream(F.device).cuda_stream
    st = (ctypes_i64 * 4)(fr, fc, ur, uc)
    rc = lib.chp_frontal_add_update(F.data_ptr(), N, d_r.data_ptr(), U.data_ptr(), st,
                                    float(alpha), stream)
    _check(rc)
    return None


def frontal_get_update(F, relidx, relptr, i):
    """Return U = lower(F[r, r]) as a new N x N float64 CUDA tensor
    (cbase.c:39-59); the strict upper part of U is zero."""
    import torch
    lib = _native.lib()
    r = np.asarray(relidx[int(relptr[i]):int(relptr[i + 1])], dtype=np.int64)
    N = len(r)
    U = torch.zeros((N, N), dtype=torch.float64, device=F.device)
    fr, fc = _strides_of(F)
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
