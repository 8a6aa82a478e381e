"""NumPy restatement of CHOMPACK's supernodal multifrontal numeric sweeps.

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg; the product
(chompack_amd) never imports it.  Each function restates one reference
routine supernode by supernode with dense LAPACK calls, following the file
and lines cited in its docstring, on float64 numpy copies of ``blkval``.

Parity of this restatement is pinned in tests/test_oracle.py against
(a) the reference's own C kernels built from /root/reference by
oracle/Makefile (oracle/_ref), (b) dense closed forms (numpy Cholesky of
P A P^T, dense inverse restricted to the filled pattern, dense triangular
solves), and (c) the reference test suite's round-trip identities
(tests/test_numeric.py of the reference).

Block layout (symbolic.py:629-632): supernode k owns blkval[blkptr[k]:
blkptr[k+1]], a column-major nj x nn array; ``_blk`` returns a writable
(nj, nn) view of it.
"""
import numpy as np
import scipy.linalg as sla

__all__ = ["cholesky", "llt", "projected_inverse", "trsm", "dot", "logdet"]


class _S(object):
    def __init__(self, symb):
        self.snpost = [int(x) for x in symb.snpost]
        self.snptr = np.asarray(symb.snptr)
        self.relptr = np.asarray(symb.relptr)
        self.relidx = np.asarray(symb.relidx)
        self.chptr = np.asarray(symb.chptr)
        self.chidx = np.asarray(symb.chidx)
        self.blkptr = np.asarray(symb.blkptr)
        self.p = np.asarray(symb.p)
        self.n = symb.n

    def dims(self, k):
        nn = int(self.snptr[k + 1] - self.snptr[k])
        na = int(self.relptr[k + 1] - self.relptr[k])
        return nn, na, nn + na

    def rel(self, c):
        return self.relidx[self.relptr[c]:self.relptr[c + 1]]

    def children(self, k):
        return [int(c) for c in self.chidx[self.chptr[k]:self.chptr[k + 1]]]


def _blk(blkval, S, k):
    nn, na, nj = S.dims(k)
    return blkval[S.blkptr[k]:S.blkptr[k + 1]].reshape((nn, nj)).T


def _sym_lower(M):
    """Symmetric matrix from the lower triangle of M."""
    return np.tril(M) + np.tril(M, -1).T


def _write_lower(dst, src, nn):
    """dlacpy('L') of an nj x nn block: the lower trapezoid only."""
    nj = dst.shape[0]
    ii, jj = np.tril_indices(nj, 0, nn)
    dst[ii, jj] = src[ii, jj]


def _extend_add(F, U, r):
    """F[r[i], r[j]] += U[i, j] for i >= j (cbase.c:11-34, cholesky.c:41-52)."""
    N = len(r)
    ii, jj = np.tril_indices(N)
    F[r[ii], r[jj]] += U[ii, jj]


def _extract(F, r):
    """U[i, j] = F[r[i], r[j]] for i >= j, zero above (cbase.c:39-59)."""
    N = len(r)
    U = np.zeros((N, N))
    ii, jj = np.tril_indices(N)
    U[ii, jj] = F[r[ii], r[jj]]
    return U


def cholesky(symb, blkval):
    """Supernodal multifrontal Cholesky, leaves to root (cholesky.c:26-82).
    Overwrites blkval with the factor in the stored convention: the (2,1)
    block of supernode k holds L_{Ak,Nk} L_{Nk,Nk}^{-1} (cholesky.c:71-72).
    Returns 0 or the LAPACK-style info of the failing diagonal block."""
    S = _S(symb)
    stack = []
    for k in S.snpost:
        nn, na, nj = S.dims(k)
        B = _blk(blkval, S, k)
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(B, 0) if nn == nj else np.vstack([np.tril(B[:nn]), B[nn:]])
        for c in reversed(S.children(k)):
            _extend_add(F, stack.pop(), S.rel(c))
        try:
            L11 = np.linalg.cholesky(_sym_lower(F[:nn, :nn]))
        except np.linalg.LinAlgError:
            d = np.linalg.eigvalsh(_sym_lower(F[:nn, :nn]))
            return max(1, int(np.sum(d > 0)) + 1)
        F[:nn, :nn] = L11
        if na > 0:
            L21 = sla.solve_triangular(L11, F[nn:, :nn].T, lower=True).T
            F22 = np.tril(F[nn:, nn:]) - np.tril(L21 @ L21.T)
            F[nn:, :nn] = sla.solve_triangular(L11, L21.T, lower=True, trans="T").T
            stack.append(np.tril(F22))
        _write_lower(B, F[:, :nn], nn)
    return 0


def llt(symb, blkval):
    """Chordal product X = L L^T, leaves to root (llt.c:26-58).  The whole
    nj x nn block (including the strict upper part of its diagonal block, as
    dsyrk reads it) forms the product."""
    S = _S(symb)
    stack = []
    for k in S.snpost:
        nn, na, nj = S.dims(k)
        B = _blk(blkval, S, k)
        B[nn:, :] = B[nn:, :] @ np.tril(B[:nn, :])
        F = np.tril(B @ B.T)
        for c in reversed(S.children(k)):
            _extend_add(F, stack.pop(), S.rel(c))
        if na > 0:
            stack.append(np.tril(F[nn:, nn:]))
        _write_lower(B, F[:, :nn], nn)
    return 0


def projected_inverse(symb, blkval):
    """Y = P(X^{-1}) from the factor, root to leaves
    (projected_inverse.c:25-73).  Returns 0 or the dtrtri info."""
    S = _S(symb)
    stack = []
    for k in reversed(S.snpost):
        nn, na, nj = S.dims(k)
        B = _blk(blkval, S, k)
        L11 = np.tril(B[:nn, :])
        d = np.diag(L11)
        if np.any(d == 0):
            return int(np.argmax(d == 0)) + 1
        Li = sla.solve_triangular(L11, np.eye(nn), lower=True)
        B[:nn, :] = np.tril(Li)
        F = np.zeros((nj, nj))
        F[:nn, :nn] = Li.T @ Li
        if na > 0:
            V = _sym_lower(stack.pop())
            B21 = B[nn:, :]
            F21 = -V @ B21
            F[:nn, :nn] -= F21.T @ B21
            F[nn:, :nn] = F21
            F[:nn, nn:] = F21.T
            F[nn:, nn:] = V
        for c in S.children(k):
            stack.append(_extract(F, S.rel(c)))
        _write_lower(B, F[:, :nn], nn)
    return 0


def trsm(symb, lblkval, B, alpha=1.0, trans="N"):
    """B := alpha L^{-1} B ('N', leaves to root) or alpha L^{-T} B ('T',
    root to leaves) with the chordal factor L and B in the original ordering
    (trsm.c:32-136).  B is an (n, nrhs) float64 array, modified in place."""
    S = _S(symb)
    B2 = B.reshape((S.n, -1))
    stack = []
    if trans == "N":
        for k in S.snpost:
            nn, na, nj = S.dims(k)
            L = _blk(lblkval, S, k)
            rows = S.p[S.snptr[k]:S.snptr[k + 1]]
            x = np.zeros((nj, B2.shape[1]))
            x[:nn] = alpha * B2[rows]
            for c in reversed(S.children(k)):
                x[S.rel(c)] += stack.pop()
            if na > 0:
                x[nn:] -= L[nn:, :] @ x[:nn]
                stack.append(x[nn:].copy())
            x[:nn] = sla.solve_triangular(np.tril(L[:nn]), x[:nn], lower=True)
            B2[rows] = x[:nn]
    elif trans == "T":
        for k in reversed(S.snpost):
            nn, na, nj = S.dims(k)
            L = _blk(lblkval, S, k)
            rows = S.p[S.snptr[k]:S.snptr[k + 1]]
            x = np.zeros((nj, B2.shape[1]))
            x[:nn] = alpha * B2[rows]
            x[:nn] = sla.solve_triangular(np.tril(L[:nn]), x[:nn], lower=True, trans="T")
            if na > 0:
                x[nn:] = stack.pop()
                x[:nn] -= L[nn:, :].T @ x[nn:]
            for c in S.children(k):
                stack.append(x[S.rel(c)].copy())
            B2[rows] = x[:nn]
    else:
        raise ValueError("trans must be 'N' or 'T'")


def dot(symb, x, y):
    """tr(XY) = 2 * sum(lower) - sum(diag) over the blocks (dot.c:3-22)."""
    S = _S(symb)
    val = 0.0
    for k in range(symb.Nsn):
        nn, na, nj = S.dims(k)
        X = _blk(x, S, k)
        Y = _blk(y, S, k)
        for j in range(nn):
            val -= X[j, j] * Y[j, j]
            val += 2.0 * float(np.dot(X[j:, j], Y[j:, j]))
    return val


def logdet(symb, lblkval):
    """log det X = 2 sum log diag(L11) from a factor (symbolic.py:1113-1128)."""
    S = _S(symb)
    s = 0.0
    for k in range(symb.Nsn):
        s += float(np.sum(np.log(np.diag(_blk(lblkval, S, k)[:S.dims(k)[0]]))))
    return 2.0 * s
