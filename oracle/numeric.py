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

__all__ = ["cholesky", "cholesky_adjoint", "blk_weights", "llt", "projected_inverse", "completion", "update_factor", "hessian",
           "trsm", "trmm", "pf_factor", "pf_solve", "pf_product", "dot", "logdet"]


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


def cholesky_adjoint(symb, lblkval, gblkval):
    """Reverse-mode sweep of the multifrontal Cholesky (cholesky.c:26-82 run
    backwards, root to leaves).  lblkval: the factor in the stored convention
    ([L11; B21] per supernode, B21 = L21 L11^{-1}); gblkval: the gradient of a
    scalar loss with respect to that stored factor (lower trapezoid read).
    Returns fbar (blkval-shaped): per supernode the lower trapezoid of the
    symmetric gradient of its assembled frontal matrix, so that the gradient
    with respect to the lower-stored entries of X is fbar * w with w = 2 on
    strictly-lower entries and 1 on the diagonal.  Per front, with X11 =
    L11^{-1}, L21 = B21 L11 and Ub the symmetric gradient of its update
    matrix (from the parent; 0 at a root):
        T = Gb21 X11^T,  Lb21 = T - 2 Ub L21,  G21 = Lb21 X11,
        Lb = tril(Gb11 - B21^T T - G21^T L21),  P = Phi(L11^T Lb),
        S11 = sym(X11^T P X11),  Fbar = [S11, G21^T / 2; G21 / 2, Ub],
    Phi = lower triangle with the diagonal halved; each child c receives
    Ub_c = Fbar[r_c, r_c]."""
    S = _S(symb)
    out = np.zeros_like(lblkval)
    ubar = {}
    for k in reversed(S.snpost):
        nn, na, nj = S.dims(k)
        B = _blk(lblkval, S, k)
        G = _blk(gblkval, S, k)
        L11 = np.tril(B[:nn, :])
        X11 = sla.solve_triangular(L11, np.eye(nn), lower=True)
        B21 = B[nn:, :]
        L21 = B21 @ L11
        Ub = ubar.pop(k, np.zeros((na, na)))
        T = G[nn:, :] @ X11.T
        Lb21 = T - 2.0 * Ub @ L21
        G21 = Lb21 @ X11
        Lb = np.tril(np.tril(G[:nn, :]) - B21.T @ T - G21.T @ L21)
        P = np.tril(L11.T @ Lb)
        P[np.diag_indices(nn)] *= 0.5
        Sm = X11.T @ P @ X11
        F = np.zeros((nj, nj))
        F[:nn, :nn] = 0.5 * (Sm + Sm.T)
        F[nn:, :nn] = 0.5 * G21
        F[:nn, nn:] = 0.5 * G21.T
        F[nn:, nn:] = Ub
        for c in S.children(k):
            r = S.rel(c)
            ubar[c] = F[np.ix_(r, r)]
        _write_lower(_blk(out, S, k), F[:, :nn], nn)
    return out


def blk_weights(symb):
    """w of cholesky_adjoint: 2 on strictly-lower entries of every block,
    1 on the diagonal, 0 above it (the gradient with respect to a lower-stored
    entry x_ij counts both x_ij and x_ji)."""
    S = _S(symb)
    w = np.zeros(int(S.blkptr[-1]))
    for k in range(len(S.snptr) - 1):
        nn, na, nj = S.dims(k)
        Wk = _blk(w, S, k)
        ii, jj = np.tril_indices(nj, -1, nn)
        Wk[ii, jj] = 2.0
        Wk[np.arange(nn), np.arange(nn)] = 1.0
    return w


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


def trmm(symb, lblkval, B, alpha=1.0, trans="N"):
    """B := alpha L B ('N', leaves to root) or alpha L^T B ('T', root to
    leaves) with the chordal factor L and B in the original ordering
    (pybase/trmm.py:44-114).  With the stored B21 = L21 L11^{-1}, 'N' forms
    [L11 b; B21 (L11 b)] and then adds the children's contributions (popped,
    i.e. in reverse child order); 'T' receives the parent's pre-product rows,
    adds B21^T u_A and multiplies by L11^T.  B is an (n, nrhs) float64 array,
    modified in place."""
    S = _S(symb)
    B2 = B.reshape((S.n, -1))
    stack = []
    if trans == "N":
        for k in S.snpost:
            nn, na, nj = S.dims(k)
            L = _blk(lblkval, S, k)
            rows = S.p[S.snptr[k]:S.snptr[k + 1]]
            u = np.zeros((nj, B2.shape[1]))
            u[:nn] = np.tril(L[:nn]) @ (alpha * B2[rows])
            if na > 0:
                u[nn:] = L[nn:, :] @ u[:nn]
            for c in reversed(S.children(k)):
                u[S.rel(c)] += stack.pop()
            if na > 0:
                stack.append(u[nn:].copy())
            B2[rows] = u[:nn]
    elif trans == "T":
        for k in reversed(S.snpost):
            nn, na, nj = S.dims(k)
            L = _blk(lblkval, S, k)
            rows = S.p[S.snptr[k]:S.snptr[k + 1]]
            u = np.zeros((nj, B2.shape[1]))
            u[:nn] = alpha * B2[rows]
            if na > 0:
                u[nn:] = stack.pop()
            for c in S.children(k):
                stack.append(u[S.rel(c)].copy())
            x = u[:nn].copy()
            if na > 0:
                x += L[nn:, :].T @ u[nn:]
            B2[rows] = np.tril(L[:nn]).T @ x
    else:
        raise ValueError("trans must be 'N' or 'T'")


def _ddrsv(trans, v, l, b, x):
    """x := L^{-1} x ('N') or L^{-T} x ('T'), L = diag(l) + strictly-lower
    (v b^T) (pfcholesky.c:4-25)."""
    s = 0.0
    n = len(x)
    if trans == "N":
        for k in range(n):
            x[k] = (x[k] - v[k] * s) / l[k]
            s += x[k] * b[k]
    else:
        for k in range(n - 1, -1, -1):
            x[k] = (x[k] - b[k] * s) / l[k]
            s += x[k] * v[k]


def _ddrmv(trans, v, l, b, x):
    """x := L x ('N') or L^T x ('T') (pfcholesky.c:27-50)."""
    s = 0.0
    n = len(x)
    if trans == "N":
        for k in range(n):
            tmp = x[k] * b[k]
            x[k] = l[k] * x[k] + v[k] * s
            s += tmp
    else:
        for k in range(n - 1, -1, -1):
            tmp = x[k] * v[k]
            x[k] = l[k] * x[k] + b[k] * s
            s += tmp


def pf_factor(a, V, L, B):
    """Product-form factorization of I + V diag(a) V^T in place
    (dpftrf, pfcholesky.c:52-87); returns 0, or -1 where the reference
    fails."""
    n, m = V.shape
    for i in range(m):
        v, l, b = V[:, i], L[:, i], B[:, i]
        alpha = -float(a[i])
        for j in range(n):
            tmp = 1.0 - alpha * v[j] * v[j]
            if tmp <= 0:
                return -1
            l[j] = np.sqrt(tmp)
            b[j] = -alpha * v[j] / l[j]
            alpha += b[j] * b[j]
        for j in range(i + 1, m):
            _ddrsv("N", v, l, b, V[:, j])
    return 0


def pf_solve(V, L, B, U, trans="N"):
    """U := (L_1 ... L_m)^{-1} U or its transpose, column by column
    (dpfsv, pfcholesky.c:89-112)."""
    m = V.shape[1]
    order = range(m) if trans == "N" else range(m - 1, -1, -1)
    for c in range(U.shape[1]):
        for i in order:
            _ddrsv(trans, V[:, i], L[:, i], B[:, i], U[:, c])


def pf_product(V, L, B, U, trans="N"):
    """U := L_1 ... L_m U or its transpose (dpfmv, pfcholesky.c:114-136)."""
    m = V.shape[1]
    order = range(m - 1, -1, -1) if trans == "N" else range(m)
    for c in range(U.shape[1]):
        for i in order:
            _ddrmv(trans, V[:, i], L[:, i], B[:, i], U[:, c])


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


PIVOT_NOISE = 1e-13


def _larfg(alpha, x):
    """Householder reflector H = I - tau [1; v][1; v]^T with
    H [alpha; x] = [beta; 0], beta = -sign(alpha) ||[alpha; x]|| (LAPACK dlarfg).
    A pivot with |alpha| <= PIVOT_NOISE * ||x|| is taken as +0 (beta = -||x||),
    the exact-arithmetic value of pivots that vanish structurally: their
    computed sign is rounding noise, which would otherwise decide the sign of
    a row of the factor (see update_factor).  Returns (tau, beta, v)."""
    xnorm = float(np.linalg.norm(x))
    if xnorm == 0.0:
        return 0.0, alpha, x.copy()
    if abs(alpha) <= PIVOT_NOISE * xnorm:
        alpha = 0.0
    beta = -np.copysign(np.hypot(alpha, xnorm), alpha)
    tau = (beta - alpha) / beta
    return tau, beta, x * (1.0 / (alpha - beta))


def _rev_chol_inverse_factor(S):
    """Lower-triangular Lr with S = Lr^T Lr: Cholesky of the reversed matrix
    J S J = U^T U, then Lr = J U J (completion.c:93-104, update_factor.c:67-103).
    Raises LinAlgError if S is not positive definite."""
    U = np.linalg.cholesky(S[::-1, ::-1]).T
    return np.ascontiguousarray(U[::-1, ::-1])


def update_factor(r, nn, na, F):
    """Factor of the child update V = F[r, r] = Li^T Li, Li lower triangular,
    when the (2,2) block of the frontal matrix F holds the parent's factor Lk
    (V_k = Lk^T Lk) instead of V_k (update_factor.c:3-108, misc.py:84-176).
    The separator part C = Lk[:, r2 - nn] is reduced to lower-triangular form
    by Householder reflections from the last column, Li21 solves
    Li22^T Li21 = B, and Li11 is the reversed Cholesky factor of
    A - Li21^T Li21.  Returns Li (N x N, zero above the diagonal).

    Li is unique up to the signs of the rows of Li22.  Li^T Li (all that the
    completion and the full Hessian map use) does not depend on them, but the
    Hessian factors G_X, G_X^adj with factored updates do; the reflector
    pivots can vanish structurally, and then the reference's signs follow
    the rounding of its BLAS.  Here such pivots count as +0 (_larfg)."""
    r = np.asarray(r)
    N = len(r)
    N1 = int(np.sum(r < nn))
    N2 = N - N1
    Li = np.zeros((N, N))
    Li[:, :N1] = F[np.ix_(r, r[:N1])]
    if N2 > 0:
        C = np.zeros((na, N2))
        for j in range(N2):
            jj = int(r[N1 + j])
            C[jj - nn:, j] = F[jj:, jj]
        for j in range(N2):
            col = N2 - 1 - j
            jj = int(r[N - 1 - j]) - nn   # first structural row of this column
            t = (na - jj) - (j + 1)       # entries to annihilate above the pivot
            if t == 0:
                continue
            piv = na - 1 - j
            tau, beta, v = _larfg(C[piv, col], C[jj:piv, col])
            C[jj:piv, col] = v
            C[piv, col] = beta
            if col > 0 and tau != 0.0:
                w = np.concatenate([v, [1.0]])
                blk = C[jj:piv + 1, :col]
                blk -= tau * np.outer(w, w @ blk)
        L22 = np.tril(C[na - N2:, :])
        Li[N1:, N1:] = L22
        L21 = sla.solve_triangular(L22, Li[N1:, :N1], lower=True, trans="T")
        Li[N1:, :N1] = L21
        A = np.tril(Li[:N1, :N1]) - np.tril(L21.T @ L21)
    else:
        A = np.tril(Li[:N1, :N1])
    Li[:N1, :N1] = _rev_chol_inverse_factor(_sym_lower(A))
    return Li


def completion(symb, blkval, factored_updates=True):
    """Maximum-determinant positive definite completion, root to leaves
    (completion.c:3-120): on exit blkval holds the factor L (stored
    convention) with P((L L^T)^{-1}) = X.  Returns 0 or a LAPACK-style info
    (the caller raises and still marks the matrix a factor, cbase.c:612-617)."""
    S = _S(symb)
    stack = []
    for k in reversed(S.snpost):
        nn, na, nj = S.dims(k)
        B = _blk(blkval, S, k)
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(B, 0) if nn == nj else np.vstack([np.tril(B[:nn]), B[nn:]])
        if na > 0:
            F[nn:, nn:] = stack.pop()
        try:
            for c in S.children(k):
                if factored_updates:
                    stack.append(update_factor(S.rel(c), nn, na, F))
                else:
                    stack.append(_extract(F, S.rel(c)))
            if na > 0:
                X21 = B[nn:, :].copy()
                if factored_updates:      # V_k = Lk^T Lk
                    Lk = np.tril(F[nn:, nn:])
                    Z = sla.solve_triangular(Lk, X21, lower=True, trans="T")
                    W = lambda Z: sla.solve_triangular(Lk, Z, lower=True)
                else:                     # V_k = Lk Lk^T
                    Lk = np.linalg.cholesky(_sym_lower(F[nn:, nn:]))
                    Z = sla.solve_triangular(Lk, X21, lower=True)
                    W = lambda Z: sla.solve_triangular(Lk, Z, lower=True, trans="T")
                B[:nn, :] = np.tril(B[:nn, :]) - np.tril(Z.T @ Z) + np.triu(B[:nn, :], 1)
                B[nn:, :] = -W(Z)
            Lr = _rev_chol_inverse_factor(_sym_lower(B[:nn, :]))
        except np.linalg.LinAlgError:
            return 1
        Linv = sla.solve_triangular(Lr, np.eye(nn), lower=True)
        iu = np.triu_indices(nn, 1)
        keep = B[:nn, :][iu]
        B[:nn, :] = np.tril(Linv)
        B[:nn, :][iu] = keep
    return 0


def _y2k(S, lblkval, ub, inv):
    """Leaves-to-root half of the Hessian factor G (hessian.c:3-101)."""
    alpha = 1.0 if inv else -1.0
    stack = []
    for k in S.snpost:
        nn, na, nj = S.dims(k)
        L = _blk(lblkval, S, k)
        U = _blk(ub, S, k)
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(U, 0) if nn == nj else np.vstack([np.tril(U[:nn]), U[nn:]])
        def ext_add():
            for c in reversed(S.children(k)):
                _extend_add(F, stack.pop(), S.rel(c))
        if not inv:
            ext_add()
        if na > 0:
            L21 = L[nn:, :]
            F[nn:, nn:] += alpha * (L21 @ F[nn:, :nn].T)
            F[nn:, :nn] += alpha * (L21 @ _sym_lower(F[:nn, :nn]))
            F[nn:, nn:] += alpha * (F[nn:, :nn] @ L21.T)
        if inv:
            ext_add()
        if na > 0:
            stack.append(np.tril(F[nn:, nn:]))
        _write_lower(U, F[:, :nn], nn)


def _m2t(S, lblkval, ub, inv):
    """Root-to-leaves half of the Hessian factor (hessian.c:103-198)."""
    alpha = 1.0 if inv else -1.0
    stack = []
    for k in reversed(S.snpost):
        nn, na, nj = S.dims(k)
        L = _blk(lblkval, S, k)
        U = _blk(ub, S, k)
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(U, 0) if nn == nj else np.vstack([np.tril(U[:nn]), U[nn:]])
        if na > 0:
            F[nn:, nn:] = _sym_lower(stack.pop())
        if inv:
            for c in S.children(k):
                stack.append(_extract(F, S.rel(c)))
        if na > 0:
            L21 = L[nn:, :]
            F[:nn, :nn] += alpha * (F[nn:, :nn].T @ L21)
            F[nn:, :nn] += alpha * (_sym_lower(F[nn:, nn:]) @ L21)
            F[:nn, :nn] += alpha * (L21.T @ F[nn:, :nn])
        _write_lower(U, F[:, :nn], nn)
        if not inv:
            for c in S.children(k):
                stack.append(_extract(F, S.rel(c)))


def _scale(S, lblkval, yblkval, ubs, inv, adj, factored_updates):
    """Diagonal scaling part of the Hessian factor, root to leaves
    (hessian.c:201-352).  Returns 0 or the info of a failed update
    factorization (which the reference's hessian() discards)."""
    stack = []
    tr = lambda A, t: A.T if t == "T" else A
    for k in reversed(S.snpost):
        nn, na, nj = S.dims(k)
        L = _blk(lblkval, S, k)
        Y = _blk(yblkval, S, k)
        F = np.zeros((nj, nj))
        F[:, :nn] = np.tril(Y, 0) if nn == nj else np.vstack([np.tril(Y[:nn]), Y[nn:]])
        if na > 0:
            F[nn:, nn:] = stack.pop()
        try:
            for c in S.children(k):
                if factored_updates:
                    stack.append(update_factor(S.rel(c), nn, na, F))
                else:
                    stack.append(_extract(F, S.rel(c)))
            Lk = None
            if na > 0:
                if factored_updates:
                    Lk = np.tril(F[nn:, nn:])
                    tr3 = "T" if adj else "N"
                else:
                    Lk = np.linalg.cholesky(_sym_lower(F[nn:, nn:]))
                    tr3 = "N" if adj else "T"
        except np.linalg.LinAlgError:
            return 1
        tr1, tr2 = ("N", "T") if adj else ("T", "N")
        L11 = np.tril(L[:nn, :])
        for ub in ubs:
            U = _blk(ub, S, k)
            U[:nn, :] = _sym_lower(U[:nn, :])
            if not inv:
                # dtrsm R: U := U op1(L11)^{-1};  dtrsm L: U11 := op2(L11)^{-1} U11
                U[:, :] = sla.solve_triangular(tr(L11, tr1).T, U.T,
                                               lower=(tr1 == "T")).T
                U[:nn, :] = sla.solve_triangular(L11, U[:nn, :], lower=True,
                                                 trans=tr2)
                U[:nn, :] = np.tril(U[:nn, :])
                if na > 0:
                    U[nn:, :] = tr(Lk, tr3) @ U[nn:, :]
            else:
                U[:, :] = U @ tr(L11, tr1)
                U[:nn, :] = tr(L11, tr2) @ U[:nn, :]
                U[:nn, :] = np.tril(U[:nn, :])
                if na > 0:
                    U[nn:, :] = sla.solve_triangular(Lk, U[nn:, :], lower=True, trans=tr3)
    return 0


def hessian(symb, lblkval, yblkval, ublkvals, adj=False, inv=False, factored_updates=True):
    """Hessian factor G_X / its adjoint / their inverses on each array of
    ``ublkvals`` in place (hessian.c:354-374); ``adj=None`` applies the full
    map G^adj G (or its inverse), as the reference's wrapper does
    (cbase.c:701-713, 833-842).  L is the factor, Y the projected inverse."""
    S = _S(symb)
    inv = bool(inv)
    a0 = inv if adj is None else bool(adj)
    passes = [a0] if adj is not None else [a0, not a0]
    for a in passes:
        if a != inv:
            _scale(S, lblkval, yblkval, ublkvals, inv, a, factored_updates)
        for ub in ublkvals:
            (_m2t if a else _y2k)(S, lblkval, ub, inv)
        if a == inv:
            _scale(S, lblkval, yblkval, ublkvals, inv, a, factored_updates)
    return 0
