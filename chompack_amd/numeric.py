"""Numeric sweeps on the GPU: the drop-in for chompack.cbase's
``cholesky, llt, projected_inverse, completion, hessian, trsm`` and
``base.dot`` (src/python/__init__.py:27-33, src/C/cbase.c:106-1077).

Every function has the reference's signature, in-place semantics,
``is_factor`` state machine and exception types/messages; the work runs as
hand-written HIP kernels through the C-ABI (include/chompack_amd.h).  There is
no CPU fallback: without the HIP library or a GPU these functions raise.
"""
import ctypes

import numpy as np

from . import _native
from .symbolic import cspmatrix

__all__ = ["Plan", "cholesky", "cholesky_adjoint", "llt", "projected_inverse", "completion", "hessian", "trmm",
           "trsm", "dot", "logdet", "logdiag", "copy_blkval"]

SMALL_MAX = 64  # largest clique order handled by the LDS-resident front kernels


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("chompack_amd numeric routines need a ROCm GPU (none visible)")
    return torch


def _check_blkval(X):
    torch = _torch()
    b = X.blkval
    if not (isinstance(b, torch.Tensor) and b.is_cuda and b.dtype == torch.float64
            and b.is_contiguous()):
        raise TypeError("blkval must be a contiguous float64 tensor in GPU memory")
    return b


def _stream(t):
    torch = _torch()
    return torch.cuda.current_stream(t.device).cuda_stream


def _raise_rc(rc, what):
    if rc < 0:
        msg = _native.lib().chp_last_error()
        raise RuntimeError("%s: %s" % (what, msg.decode() if msg else "error %d" % rc))


class Plan(object):
    """Device plan of one symbolic factorization (level schedule, index arrays
    and work arrays resident in HBM).  Built on first use, cached on the
    symbolic object, freed with it."""

    PARAMS = ("small_max", "panel_width", "splitk_max", "lds_cap", "chol_dag")

    def __init__(self, symb, device, small_max=None, **params):
        """``params``: panel_width (outer panel of large fronts, a multiple
        of 64), splitk_max (1 = no split-K), lds_cap (bytes of LDS a
        small-front workgroup may use), chol_dag (large fronts of the FP64
        cholesky: 1 one tile-DAG launch per level, 2 panel and GEMM
        launches, the default); omitted ones keep their defaults."""
        torch = _torch()
        lib = _native.lib()
        self.symb_arrays = [_native.as_i64(a) for a in (
            symb.snpost, symb.snptr, symb.relptr, symb.relidx, symb.chptr, symb.chidx,
            symb.blkptr, symb.p)]
        a = self.symb_arrays
        self.device = torch.device(device)
        unknown = set(params) - set(self.PARAMS[1:])
        if unknown:
            raise TypeError("unknown plan parameters: %s" % ", ".join(sorted(unknown)))
        pv = np.array([int(small_max or SMALL_MAX)] +
                      [int(params.get(k) or 0) for k in self.PARAMS[1:]], dtype=np.int32)
        with torch.cuda.device(self.device):
            h = lib.chp_plan_create_params(symb.n, symb.Nsn, *[_native.ptr(x) for x in a],
                                           pv.ctypes.data, len(pv))
        if not h:
            raise RuntimeError("plan creation failed: %s" % lib.chp_last_error().decode())
        self.handle = h
        self._lib = lib

    def params(self):
        """The plan's parameters: small_max, panel_width, splitk_max, lds_cap, chol_dag."""
        out = np.zeros(5, dtype=np.int32)
        self._lib.chp_plan_params(self.handle, out.ctypes.data)
        return dict(zip(self.PARAMS, [int(v) for v in out]))

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        self._lib.chp_plan_info(self.handle, _native.ptr(out))
        keys = ["levels", "large_chol", "large_projinv", "gemm_tiles", "upd_doubles",
                "dinv_doubles", "scratch_doubles", "chol_launches"]
        return dict(zip(keys, [int(v) for v in out]))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                self._lib.chp_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _plan(X):
    b = _check_blkval(X)
    return X.symb._plan(b.device), b


def cholesky(X):
    """Supernodal multifrontal Cholesky factorization X = L L^T, in place
    (cbase.c:187-284).  On exit X holds the factor (stored convention: the
    (2,1) block of supernode k is L_{Ak,Nk} L_{Nk,Nk}^{-1})."""
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    plan, b = _plan(X)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_cholesky(plan.handle, b.data_ptr(), _stream(b), _native.ptr(info))
    X.is_factor = True
    _raise_rc(rc, "cholesky")
    if rc == 1:
        raise ArithmeticError("factorization failed")


def llt(L):
    """Chordal product X = L L^T, in place (cbase.c:298-390)."""
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("X must be a cspmatrix")
    plan, b = _plan(L)
    rc = _native.lib().chp_llt(plan.handle, b.data_ptr(), _stream(b))
    _raise_rc(rc, "llt")
    L.is_factor = False


def projected_inverse(L):
    """Projected inverse Y = P(L^{-T} L^{-1}), in place (cbase.c:404-500)."""
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("X must be a cspmatrix")
    plan, b = _plan(L)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_projected_inverse(plan.handle, b.data_ptr(), _stream(b),
                                             _native.ptr(info))
    L.is_factor = False
    _raise_rc(rc, "projected_inverse")
    if rc == 1:
        raise ArithmeticError("partial inverse failed")


def cholesky_adjoint(L, G):
    """Reverse-mode sweep of ``cholesky`` (cholesky.c:26-82 run backwards):
    L is a factor (stored convention), G a cspmatrix on the same symbolic
    object holding the gradient of a scalar loss with respect to L's stored
    entries (lower trapezoid of each block).  Returns a new cspmatrix Fbar:
    per supernode the lower trapezoid of the symmetric gradient of its
    frontal matrix; the gradient with respect to the lower-stored entries of
    the factored matrix is Fbar with its strictly-lower entries doubled.
    Raises ArithmeticError on a zero pivot of L."""
    torch = _torch()
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("L must be a cspmatrix factor")
    if not isinstance(G, cspmatrix) or G.symb is not L.symb:
        raise ValueError("symbolic factorizations must be the same")
    plan, lb = _plan(L)
    gb = _check_blkval(G)
    if gb.device != lb.device:
        raise ValueError("all blkval arrays must be on the same device")
    out = torch.empty_like(lb)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_cholesky_adjoint(plan.handle, lb.data_ptr(), gb.data_ptr(),
                                            out.data_ptr(), _stream(lb), _native.ptr(info))
    _raise_rc(rc, "cholesky_adjoint")
    if rc == 1:
        raise ArithmeticError("factor has a zero pivot")
    return cspmatrix(L.symb, blkval=out)


def completion(X, factored_updates=True):
    """Maximum-determinant positive definite completion, in place
    (cbase.c:521-621, completion.c:3-120): on exit X holds the Cholesky
    factor L of the inverse of the completion, P((L L^T)^{-1}) = X.
    ``factored_updates`` selects the Householder-factored update matrices
    (update_factor.c) or plain extraction plus a Cholesky per front.  As in
    the reference, X is marked a factor even when the completion fails."""
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    plan, b = _plan(X)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_completion(plan.handle, b.data_ptr(), 1 if factored_updates else 0,
                                      _stream(b), _native.ptr(info))
    X.is_factor = True
    _raise_rc(rc, "completion")
    if rc == 1:
        raise ArithmeticError("completion failed")


def hessian(L, Y, U, adj=False, inv=False, factored_updates=True):
    """Factors of the Hessian of the log-det barrier, applied in place to U
    (a cspmatrix or a list of them) (cbase.c:675-857, hessian.c:3-374):

    ======  =====  ==============================
    adj     inv    U :=
    ======  =====  ==============================
    False   False  G_X(U)
    True    False  G_X^adj(U)
    False   True   G_X^{-1}(U)
    True    True   (G_X^adj)^{-1}(U)
    None    False  G_X^adj(G_X(U)) = P(X^{-1} U X^{-1})
    None    True   the inverse of the above
    ======  =====  ==============================

    L is the factor of X and Y = P(X^{-1}); all arguments must share one
    symbolic object.  ``adj`` and ``inv`` default to False (the reference's
    C wrapper leaves them uninitialized when omitted, cbase.c:701-713)."""
    torch = _torch()
    syms = [Y.symb] if isinstance(Y, cspmatrix) else [None]
    if isinstance(U, list):
        Us = U
    else:
        Us = [U]
    syms += [getattr(u, "symb", None) for u in Us]
    if any(s is not L.symb for s in syms):
        raise ValueError("symbolic factorizations must be the same")
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("L must be a cspmatrix factor")
    if not (isinstance(Y, cspmatrix) and Y.is_factor is False):
        raise ValueError("Y must be a cspmatrix")
    plan, lb = _plan(L)
    yb = _check_blkval(Y)
    ubs = [_check_blkval(u) for u in Us]
    for t in [yb] + ubs:
        if t.device != lb.device:
            raise ValueError("all blkval arrays must be on the same device")
    arr = (ctypes.c_void_p * max(len(ubs), 1))(*[u.data_ptr() for u in ubs])
    a = -1 if adj is None else (1 if adj else 0)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_hessian(plan.handle, lb.data_ptr(), yb.data_ptr(),
                                   ctypes.cast(arr, ctypes.c_void_p), len(ubs), a,
                                   1 if inv else 0, 1 if factored_updates else 0,
                                   _stream(lb), _native.ptr(info))
    _raise_rc(rc, "hessian")
    del torch


def colmajor_work(T):
    """(work, ld, avail, copied) for a float64 GPU vector or matrix T.

    ``work`` addresses T's elements column-major with leading dimension
    ``ld`` from ``work.data_ptr()``: T itself when its strides allow that
    (unit row stride, column stride >= rows), otherwise a column-major copy
    (``copied``; the caller writes it back with ``T.copy_(work)``).
    ``avail`` is the number of float64 elements of storage from
    ``work.data_ptr()`` on, the bound for explicit ld / column counts."""
    if T.dim() == 1:
        rows, cols = T.shape[0], 1
        ok = T.stride(0) == 1 or rows <= 1
        ld = max(rows, 1)
    elif T.dim() == 2:
        rows, cols = T.shape
        ok = (T.stride(0) == 1 or rows <= 1) and (cols <= 1 or T.stride(1) >= max(rows, 1))
        ld = T.stride(1) if (ok and cols > 1) else max(rows, 1)
    else:
        raise TypeError("expected a vector or a matrix")
    if ok:
        avail = T.untyped_storage().nbytes() // 8 - T.storage_offset()
        return T, ld, avail, False
    import torch
    work = torch.empty_strided(tuple(T.shape), (1,) if T.dim() == 1 else (1, max(rows, 1)),
                               dtype=T.dtype, device=T.device)
    work.copy_(T)
    return work, max(rows, 1), max(rows, 1) * cols, True


def host_dense(B):
    """Writable float64 numpy view of a host dense matrix or vector (numpy
    array, or an object exposing the buffer protocol / array interface such
    as a cvxopt ``matrix``, whose storage is column-major), sharing its
    memory, so an in-place routine can write its result back."""
    a = np.asarray(B)
    if a.dtype != np.float64:
        raise TypeError("B must hold float64 values")
    if a.ndim not in (1, 2):
        raise TypeError("B must be a vector or a matrix")
    if not a.flags.writeable:
        raise ValueError("B must be writable (it is updated in place)")
    return a


def _triangular(fn, what, L, B, alpha, trans, nrhs, offsetB, ldB):
    """Shared argument handling of trsm and trmm (cbase.c:965-1041).

    B is a float64 GPU tensor (any strides: storage that is not column-major
    is solved through a column-major copy written back afterwards) or a
    host dense matrix (numpy array or a buffer-protocol object such as a
    cvxopt ``matrix``), which is staged through the GPU and overwritten in
    place, as the reference overwrites its ``matrix`` argument."""
    torch = _torch()
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("L must be a cspmatrix factor")
    if trans not in ('N', 'T'):
        raise ValueError("trans must be 'N' or 'T'")
    plan, b = _plan(L)
    n = L.symb.n
    host = None
    if isinstance(B, torch.Tensor):
        if not (B.is_cuda and B.dtype == torch.float64):
            raise TypeError("B must be a float64 tensor in GPU memory or a host matrix")
        if B.device != b.device:
            raise ValueError("B must be on the device of L")
        if B.dim() not in (1, 2):
            raise TypeError("B must be a vector or a matrix")
        work, ld, avail, copied = colmajor_work(B)
        cols = 1 if B.dim() == 1 else B.shape[1]
    else:
        host = host_dense(B)
        # the column-major element sequence of the host matrix (its own
        # storage for a Fortran-ordered buffer such as cvxopt's)
        flat = host.reshape(-1, order="F")
        work = torch.as_tensor(np.ascontiguousarray(flat)).to(b.device)
        ld = max(host.shape[0], 1)
        avail = work.numel()
        copied = False
        cols = 1 if host.ndim == 1 else host.shape[1]
    if ldB != -1 and ldB is not None:
        ld = int(ldB)
    k = cols if (nrhs == -1 or nrhs is None) else int(nrhs)
    if int(offsetB) < 0:
        raise ValueError("offsetB must be non-negative")
    if ld < max(n, 1):
        raise ValueError("ldB must be at least n")
    if k < 0:
        raise ValueError("nrhs must be non-negative")
    if k > 0 and int(offsetB) + (k - 1) * ld + n > avail:
        raise ValueError("B is too small for n, nrhs, offsetB and ldB")
    rc = fn(plan.handle, b.data_ptr(), work.data_ptr() + 8 * int(offsetB),
            int(ld), int(k), float(alpha), ord(trans), _stream(b))
    _raise_rc(rc, what)
    if copied:
        B.copy_(work)
    if host is not None:
        res = work.cpu().numpy()
        if host.ndim == 1:
            host[:] = res
        else:
            host[...] = res.reshape(host.shape, order="F")


def trsm(L, B, alpha=1.0, trans='N', nrhs=-1, offsetB=0, ldB=-1):
    """B := alpha L^{-1} B (trans 'N') or alpha L^{-T} B ('T') with the chordal
    factor L, B in the original ordering (cbase.c:965-1077).

    B is a float64 CUDA tensor (a vector of length n, or an (n, k) matrix)
    or a host dense matrix (numpy, or a cvxopt-style ``matrix`` through the
    buffer protocol), updated in place.  As in the reference, ``nrhs``,
    ``offsetB`` and ``ldB`` address B's column-major storage; a tensor whose
    strides are not column-major is solved through a column-major copy that
    is written back."""
    _triangular(_native.lib().chp_trsm, "trsm", L, B, alpha, trans, nrhs, offsetB, ldB)


def trmm(L, B, alpha=1.0, trans='N', nrhs=None, offsetB=0, ldB=None):
    """B := alpha L B (trans 'N') or alpha L^T B ('T') with the chordal factor
    L, B in the original ordering (pybase/trmm.py:4-114; the reference's
    defaults nrhs = None and ldB = None mean B's column count and row
    count).  B as for ``trsm``."""
    _triangular(_native.lib().chp_trmm, "trmm", L, B, alpha, trans, nrhs, offsetB, ldB)


def dot(X, Y):
    """tr(XY) of two chordal matrices on the same symbolic object; returns
    None when the symbolic objects differ (cbase.c:106-174)."""
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    if not (isinstance(Y, cspmatrix) and Y.is_factor is False):
        raise ValueError("Y must be a cspmatrix")
    if X.symb is not Y.symb:
        return None
    plan, bx = _plan(X)
    by = _check_blkval(Y)
    out = ctypes.c_double(0.0)
    rc = _native.lib().chp_dot(plan.handle, bx.data_ptr(), by.data_ptr(), ctypes.byref(out),
                               _stream(bx))
    _raise_rc(rc, "dot")
    return out.value


def logdet(L):
    """log det X from the Cholesky factor L of X: 2 sum log diag(L), with the
    diagonal as cspmatrix.diag() reads it from a factor (symbolic.py:1113-1128)."""
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("L must be a cspmatrix factor")
    plan, b = _plan(L)
    out = ctypes.c_double(0.0)
    rc = _native.lib().chp_logdet(plan.handle, b.data_ptr(), ctypes.byref(out), _stream(b))
    _raise_rc(rc, "logdet")
    if rc == 1:
        raise ArithmeticError("factor has a non-positive diagonal entry")
    return out.value


def logdiag(L):
    """log L(j, j) for every column j of a factor, in the reordered numbering,
    as a float64 GPU tensor (log det X = 2 * logdiag(L).sum())."""
    torch = _torch()
    if not (isinstance(L, cspmatrix) and L.is_factor is True):
        raise ValueError("L must be a cspmatrix factor")
    plan, b = _plan(L)
    out = torch.empty(L.symb.n, dtype=torch.float64, device=b.device)
    rc = _native.lib().chp_logdiag(plan.handle, b.data_ptr(), out.data_ptr(), _stream(b))
    _raise_rc(rc, "logdiag")
    return out


def copy_blkval(dst, src):
    """dst[:] = src for two float64 GPU tensors of equal length (one vectorized
    device copy on the current stream)."""
    _torch()
    if dst.numel() != src.numel():
        raise ValueError("blkval arrays differ in length")
    torch = _torch()
    for t in (dst, src):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise TypeError("blkval must be a contiguous float64 GPU tensor")
    rc = _native.lib().chp_copy(dst.data_ptr(), src.data_ptr(), dst.numel(), _stream(dst))
    _raise_rc(rc, "copy")
