"""NumPy restatement of CHOMPACK's dense completions of a chordal partial
matrix.

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's
CPU-baseline leg; the product (chompack_amd) never imports it.

  psdcompletion  follows pybase/psdcompletion.py:4-70 (maximum-determinant
                 positive semidefinite completion, dense n x n);
  mrcompletion   follows pybase/mrcompletion.py:6-108 (minimum-rank
                 completion Y, n x r, with P(Y Y^T) = A);
  edmcompletion  follows pybase/edmcompletion.py:4-74 (Euclidean distance
                 matrix completion, dense n x n).

All three work on the reordered matrix exactly as the reference does (a
dense copy of P(A) for psdcompletion and edmcompletion; the clique matrices
F_k for mrcompletion),
with numpy.linalg (LAPACK) in place of cvxopt.lapack: ``cholesky`` and
``solve_triangular`` for potrf/trtrs, ``eigh`` for syevr, ``svd`` for gesvd.

Pinned in tests/test_completions.py:
  psdcompletion by (a) P(X) = A, (b) the optimality condition of the
  maximum-determinant completion (X^{-1} vanishes outside the chordal
  pattern), (c) the closed form psdcompletion(P(A^{-1})) = A^{-1} for an SPD
  A with the chordal pattern, and (d) for the pseudo-inverse branch, the
  completion of the same matrix with its zero rows and columns removed;
  mrcompletion by the reference test's own check (tests/test_mrcompletion.py
  :18-36 of the reference: P(Y Y^T) = A on its 17 x 17 example, including the
  rank-2 case built with syr2) and by r = the largest numerical rank of a
  clique.  Y itself is not unique (any Y Q with Q orthogonal also completes
  A, and LAPACK's sign choices fix one), so Y is compared through Y Y^T on
  the pattern, not entrywise;
  edmcompletion by P(X) = A, a zero diagonal and -J X J / 2 positive
  semidefinite (J the centering projection: X is a Euclidean distance
  matrix), on distances of random points.
"""
import numpy as np
import scipy.linalg as sla

__all__ = ["dense_projection", "psdcompletion", "mrcompletion", "edmcompletion", "clique_matrix"]


def dense_projection(symb, blkval):
    """Dense symmetric copy of the cspmatrix (reordered), as the reference's
    ``matrix(A.spmatrix(reordered=True, symmetric=True))``
    (psdcompletion.py:19): the lower trapezoid of every block, mirrored."""
    n = symb.n
    X = np.zeros((n, n))
    snptr, sncolptr, snrowidx, blkptr = (np.asarray(a) for a in (
        symb.snptr, symb.sncolptr, symb.snrowidx, symb.blkptr))
    bv = np.asarray(blkval, dtype=np.float64)
    for k in range(symb.Nsn):
        nn = snptr[k + 1] - snptr[k]
        rows = snrowidx[sncolptr[k]:sncolptr[k + 1]]
        nj = len(rows)
        B = bv[blkptr[k]:blkptr[k + 1]].reshape(nn, nj).T
        for j in range(nn):
            c = snptr[k] + j
            X[rows[j:], c] = B[j:, j]
            X[c, rows[j:]] = B[j:, j]
    return X


def psdcompletion(symb, blkval, reordered=True, tol=1e-15):
    """pybase/psdcompletion.py:16-70."""
    X = dense_projection(symb, blkval)
    n = symb.n
    snptr, sncolptr, snrowidx = (np.asarray(a) for a in (symb.snptr, symb.sncolptr, symb.snrowidx))
    for k in range(symb.Nsn - 1, -1, -1):                       # :28
        nn = snptr[k + 1] - snptr[k]
        beta = snrowidx[sncolptr[k]:sncolptr[k + 1]]
        nj = len(beta)
        if nj - nn == 0:                                         # :33
            continue
        alpha = beta[nn:]
        nu = beta[:nn]
        inb = np.zeros(n, dtype=bool)
        inb[beta] = True
        eta = np.nonzero(~inb[beta[0] + 1:])[0] + beta[0] + 1     # :36, indices > beta[0] not in beta
        Xaa = X[np.ix_(alpha, alpha)]
        try:                                                     # :38-48
            L = np.linalg.cholesky(Xaa)
            Xan = sla.solve_triangular(L, X[np.ix_(alpha, nu)], lower=True)
            XeaT = sla.solve_triangular(L, X[np.ix_(eta, alpha)].T, lower=True)
            tmp = XeaT.T @ Xan
        except np.linalg.LinAlgError:                            # :50-62
            w, Z = np.linalg.eigh(Xaa)
            lmax = w.max()
            d = np.array([1.0 / wi if wi > lmax * tol else 0.0 for wi in w])
            Xp = (Z * d) @ Z.T
            tmp = X[np.ix_(eta, alpha)] @ Xp @ X[np.ix_(alpha, nu)]
        X[np.ix_(eta, nu)] = tmp                                 # :64-65
        X[np.ix_(nu, eta)] = tmp.T
    if reordered:
        return X
    ip = np.asarray(symb.ip)
    return X[np.ix_(ip, ip)]


def clique_matrix(symb, blkval, k, X=None):
    """F_k = A[beta_k, beta_k] of clique k (full symmetric), the matrix that
    mrcompletion.py:135-147 assembles from blkval and the parent's V_k."""
    if X is None:
        X = dense_projection(symb, blkval)
    beta = np.asarray(symb.snrowidx)[symb.sncolptr[k]:symb.sncolptr[k + 1]]
    return X[np.ix_(beta, beta)]


def mrcompletion(symb, blkval, reordered=True):
    """pybase/mrcompletion.py:97-179.  F_k is assembled as in the reference
    (:135-147): blkval's columns of k, and V_k = the parent's F restricted to
    k's separator (frontal_get_update)."""
    n = symb.n
    snptr, sncolptr, snrowidx, blkptr, relptr, relidx, chptr, chidx = (np.asarray(a) for a in (
        symb.snptr, symb.sncolptr, symb.snrowidx, symb.blkptr, symb.relptr, symb.relidx,
        symb.chptr, symb.chidx))
    bv = np.asarray(blkval, dtype=np.float64)
    maxr = symb.clique_number
    Y = np.zeros((n, maxr))
    stack = []
    r = 0
    for k in range(symb.Nsn - 1, -1, -1):                      # :129
        nn = snptr[k + 1] - snptr[k]
        na = relptr[k + 1] - relptr[k]
        nj = na + nn
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(bv[blkptr[k]:blkptr[k + 1]].reshape(nn, nj).T)  # :137 lacpy uplo='L'
        if na > 0:
            F[nn:, nn:] = np.tril(stack.pop())                   # :140-143
        Fs = np.tril(F) + np.tril(F, -1).T
        for ii in range(chptr[k], chptr[k + 1]):                 # :146-147 frontal_get_update
            c = chidx[ii]
            ri = relidx[relptr[c]:relptr[c + 1]]
            stack.append(Fs[np.ix_(ri, ri)])
        w, Z = np.linalg.eigh(Fs)                                # :150 syevr (ascending)
        rk = int(sum(1 for wi in w[:nj] if wi > 1e-14 * w[nj - 1]))  # :151
        r = max(rk, r)
        for j in range(nj - rk, nj):                             # :155-156
            Z[:, j] *= np.sqrt(w[j])
        In = snrowidx[sncolptr[k]:sncolptr[k] + nn]
        Y[In, :rk] = Z[:nn, nj - rk:nj]                          # :158
        if na > 0:                                               # :161-174
            Ia = snrowidx[sncolptr[k] + nn:sncolptr[k + 1]]
            Ya = Y[Ia, :r].copy()
            V = np.zeros((na, r))
            V[:, :rk] = Z[nn:nj, nj - rk:nj]
            P, S1, Q1t = np.linalg.svd(V, full_matrices=True)
            _, S, Q2t = np.linalg.svd(Ya, full_matrices=True)
            Q2t = Q2t.copy()
            for i in range(min(na, rk)):
                if S[i] > 1e-14 * S[0]:
                    Q2t[i, :r] = P[:na, i] @ Y[Ia, :r] / S[i]
            Y[In, :r] = Y[In, :r] @ Q1t[:r, :r].T @ Q2t[:r, :r]
    if reordered:
        return Y[:, :r]
    return Y[np.asarray(symb.ip), :r]


def edmcompletion(symb, blkval, reordered=True, tol=1e-15):
    """pybase/edmcompletion.py:16-74."""
    X = dense_projection(symb, blkval)
    n = symb.n
    snptr, sncolptr, snrowidx = (np.asarray(a) for a in (symb.snptr, symb.sncolptr, symb.snrowidx))
    for k in range(symb.Nsn - 1, -1, -1):                       # :28
        nn = snptr[k + 1] - snptr[k]
        beta = snrowidx[sncolptr[k]:sncolptr[k + 1]]
        nj = len(beta)
        if nj - nn == 0:                                         # :33
            continue
        alpha = beta[nn:]
        nu = beta[:nn]
        inb = np.zeros(n, dtype=bool)
        inb[beta] = True
        eta = np.nonzero(~inb[beta[0] + 1:])[0] + beta[0] + 1     # :36
        a0 = alpha[0]
        h = 0.5 * X[a0, a0]
        xa, xn, xe = X[alpha, a0], X[nu, a0], X[eta, a0]
        Yaa = -0.5 * X[np.ix_(alpha, alpha)] - h + 0.5 * (xa[:, None] + xa[None, :])  # :40-41
        Ynn = -0.5 * np.diag(X)[nu] - h + xn                     # :43-44, diagonal
        Yee = -0.5 * np.diag(X)[eta] - h + xe                    # :46-47, diagonal
        Yan = -0.5 * X[np.ix_(alpha, nu)] - h + 0.5 * X[a0, nu][None, :] + 0.5 * xa[:, None]  # :49-51
        Yea = -0.5 * X[np.ix_(eta, alpha)] - h + 0.5 * X[a0, alpha][None, :] + 0.5 * xe[:, None]  # :53-55
        w, Z = np.linalg.eigh(Yaa)                               # :58-60
        lmax = w.max()
        d = np.array([1.0 / wi if wi > lmax * tol else 0.0 for wi in w])
        Yp = (Z * d) @ Z.T                                       # :63-64
        tmp = -2.0 * (Yea @ Yp @ Yan) + Ynn[None, :] + Yee[:, None]  # :67
        X[np.ix_(eta, nu)] = tmp
        X[np.ix_(nu, eta)] = tmp.T
    if reordered:
        return X
    ip = np.asarray(symb.ip)
    return X[np.ix_(ip, ip)]
