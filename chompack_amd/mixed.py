"""Mixed-precision factorization with FP64 iterative refinement.

``cholesky(X, schur)`` factors X with the Schur-complement updates of the
large fronts (the bulk of the flops) in low precision on the matrix cores --
FP32 (``schur="fp32"``) or BF16x3 split precision with FP32 accumulation
(``"bf16x3"``) -- while the diagonal blocks, the panels, the
stored-convention solves and the small LDS-resident fronts stay FP64; the
factor is accurate to about FP32 precision (FP32) or about 1e-5 relative
(BF16x3).
``solve(A, L, B, tol)`` then solves A x = b to a relative residual
max_j ||b_j - A x_j||_2 / ||b_j||_2 <= tol by iterative refinement: the
residual is formed in FP64 with the original A (``SpMat``, a device CSR copy
of the symmetric matrix) and each correction is a solve with the mixed factor.
"""
import numpy as np
import scipy.sparse as sp

from . import numeric, _native
from .misc import symmetrize
from .symbolic import cspmatrix

__all__ = ["SpMat", "cholesky", "solve"]


class SpMat(object):
    """Device CSR copy of a symmetric sparse matrix given by its lower
    triangle (scipy), in the original ordering."""

    def __init__(self, A, device="cuda:0"):
        import torch
        S = sp.csr_matrix(symmetrize(A))
        S.sort_indices()
        self.n = S.shape[0]
        self.rowptr = torch.as_tensor(S.indptr.astype(np.int64), device=device)
        self.colind = torch.as_tensor(S.indices.astype(np.int32), device=device)
        self.vals = torch.as_tensor(S.data.astype(np.float64), device=device)
        self._work = None  # row-major copy of X for several right-hand sides

    def residual(self, X, B, R):
        """R = B - A X for float64 GPU vectors or (n, k) matrices (any
        strides: operands that are not column-major go through column-major
        copies, and R's is written back)."""
        import torch
        ws = []
        for name, T in (("X", X), ("B", B), ("R", R)):
            if not (isinstance(T, torch.Tensor) and T.is_cuda and T.dtype == torch.float64):
                raise TypeError("%s must be a float64 tensor in GPU memory" % name)
            if T.dim() not in (1, 2) or T.shape[0] != self.n:
                raise ValueError("%s must have %d rows" % (name, self.n))
            ws.append(numeric.colmajor_work(T))
        k = 1 if X.dim() == 1 else X.shape[1]
        for T in (B, R):
            if (1 if T.dim() == 1 else T.shape[1]) != k:
                raise ValueError("X, B and R must have the same number of columns")
        work = None
        if k > 1:
            if self._work is None or self._work.numel() < self.n * k:
                self._work = torch.empty(self.n * k, dtype=torch.float64, device=X.device)
            work = self._work.data_ptr()
        (wx, ldx, _, _), (wb, ldb, _, _), (wr, ldr, _, rc_copy) = ws
        rc = _native.lib().chp_spmv_residual(self.n, self.rowptr.data_ptr(), self.colind.data_ptr(),
                                             self.vals.data_ptr(), wx.data_ptr(), ldx,
                                             wb.data_ptr(), ldb, wr.data_ptr(), ldr, k,
                                             work, numeric._stream(X))
        numeric._raise_rc(rc, "spmv")
        if rc_copy:
            R.copy_(wr)


SCHUR = {"fp32": 1, "bf16x3": 2}


def cholesky(X, schur="fp32"):
    """Cholesky factorization in place with low-precision Schur-complement
    updates: ``schur="fp32"`` (FP32 MFMA) or ``"bf16x3"`` (BF16x3 split
    precision with FP32 accumulation, about 2^-16 relative per product).
    Same conventions, state machine and errors as numeric.cholesky."""
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    if schur not in SCHUR:
        raise ValueError("schur must be 'fp32' or 'bf16x3'")
    plan, b = numeric._plan(X)
    info = np.zeros(2, dtype=np.int64)
    rc = _native.lib().chp_cholesky_lowprec(plan.handle, b.data_ptr(), SCHUR[schur],
                                            numeric._stream(b), _native.ptr(info))
    X.is_factor = True
    numeric._raise_rc(rc, "cholesky")
    if rc == 1:
        raise ArithmeticError("factorization failed")


def solve(A, L, B, tol=1e-10, maxiter=30):
    """x with A x = b for each column of B (column-major float64 GPU tensor,
    or a vector), from a (mixed-precision) factor L of A; A is a SpMat.
    Returns (x, residual history); the history holds max_j ||r_j|| / ||b_j||
    after each step."""
    import torch
    if B.dim() == 1:
        x = B.clone()
    else:  # a column-major copy (B.t().contiguous().t() would alias a column-major B)
        x = torch.empty_strided(B.shape, (1, B.shape[0]), dtype=B.dtype, device=B.device)
        x.copy_(B)
    numeric.trsm(L, x)
    numeric.trsm(L, x, trans="T")
    r = torch.empty_like(x)
    d = torch.empty_like(x)
    if B.dim() == 2 and B.stride(0) != 1:
        raise TypeError("B must be column-major (B.stride(0) == 1)")
    bn = torch.linalg.norm(B.reshape(B.shape[0], -1), dim=0)
    hist = []
    for it in range(maxiter + 1):
        A.residual(x, B, r)
        res = float((torch.linalg.norm(r.reshape(r.shape[0], -1), dim=0) / bn).max())
        hist.append(res)
        if res <= tol or it == maxiter:
            break
        d.copy_(r)
        numeric.trsm(L, d)
        numeric.trsm(L, d, trans="T")
        x += d
    return x, hist
