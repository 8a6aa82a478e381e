"""Pins the NumPy oracle (oracle/numeric.py) before it is trusted as the
checker of the GPU path:
  * against golden vectors written by the reference's own C kernels
    (tests/golden/*.npz, made by tests/golden/make_golden.py);
  * against oracle/_ref (the reference kernels) directly, when built;
  * against closed forms: dense Cholesky of P A P^T, the dense inverse on the
    filled pattern, dense triangular solves, dense trace;
  * against the reference suite's identities (tests/test_numeric.py:20-124).
Each check fails under a plausible oracle mistake (dropped Schur term,
missing stored-convention solve, transposed operand, wrong child order)."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import chompack_amd as cp
from chompack_amd import problems
from oracle import numeric as on
from oracle import ref

GOLD = Path(__file__).parent / "golden"
FIXTURES = sorted(p.stem for p in GOLD.glob("*.npz"))


def _problem(name):
    g = np.load(GOLD / (name + ".npz"))
    n = int(g["n"])
    A = sp.csc_matrix((g["A_vals"], (g["A_rows"], g["A_cols"])), shape=(n, n))
    p = None if name.endswith("natural") else cp.amd.order
    symb = cp.symbolic(A, p=p)
    assert np.array_equal(np.asarray(symb.p), g["symb_p"])
    assert np.array_equal(np.asarray(symb.relidx), g["symb_relidx"])
    return g, A, symb


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_vs_reference_golden(name):
    g, A, symb = _problem(name)
    b = cp.cspmatrix(symb, device="cpu").__add__(A).blkval.numpy().copy()
    assert np.array_equal(b, g["blkval_A"])
    L = b.copy()
    assert on.cholesky(symb, L) == 0
    assert _rel(L, g["blkval_L"]) < 1e-13
    Y = L.copy()
    assert on.projected_inverse(symb, Y) == 0
    assert _rel(Y, g["blkval_Y"]) < 1e-13
    P = L.copy()
    on.llt(symb, P)
    assert _rel(P, g["blkval_LLT"]) < 1e-13
    for tr in "NT":
        B = g["B"].copy()
        on.trsm(symb, g["blkval_L"].copy(), B, trans=tr)
        assert _rel(B, g["B_trsm" + tr]) < 1e-13
    assert abs(on.dot(symb, b, g["blkval_Y"]) - float(g["dot_AY"])) < 1e-10


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("side", [8, 20])
def test_oracle_vs_reference_kernels(side):
    A = problems.grid2d(side, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    b = (cp.cspmatrix(symb, device="cpu") + A).blkval.numpy().copy()
    L1, L2 = b.copy(), b.copy()
    on.cholesky(symb, L1)
    ref.cholesky(symb, L2)
    assert _rel(L1, L2) < 1e-13
    on.projected_inverse(symb, L1)
    ref.projected_inverse(symb, L2)
    assert _rel(L1, L2) < 1e-13


def _dense(A):
    return cp.symmetrize(A).toarray()


@pytest.mark.parametrize("maker", [lambda: problems.reference_test_matrix(),
                                   lambda: problems.grid2d(12, seed=0),
                                   lambda: problems.grid3d(5, seed=2)])
def test_oracle_closed_forms(maker):
    A = maker()
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cpu") + A
    b = X.blkval.numpy().copy()
    Ad = _dense(A)
    PAP = Ad[np.ix_(symb.p, symb.p)]
    Ld = np.linalg.cholesky(PAP)
    L = cp.cspmatrix(symb, blkval=b.copy(), device="cpu")
    assert on.cholesky(symb, L.blkval.numpy()) == 0
    L.is_factor = True
    Lsp = L.spmatrix(reordered=True).toarray()
    assert np.abs(Lsp - Ld).max() < 1e-12 * np.abs(Ld).max()
    # projected inverse: dense inverse on the filled pattern
    Y = L.blkval.numpy().copy()
    on.projected_inverse(symb, Y)
    Yc = cp.cspmatrix(symb, blkval=Y, device="cpu")
    Ys = Yc.spmatrix(reordered=True, symmetric=True).toarray()
    Pat = symb.sparsity_pattern(reordered=True, symmetric=True).toarray() != 0
    Ainv = np.linalg.inv(PAP)
    assert np.abs(Ys[Pat] - Ainv[Pat]).max() < 1e-12 * np.abs(Ainv).max()
    # solves: L^{-1} b and L^{-T} b in the original ordering
    rng = np.random.default_rng(1)
    B = rng.standard_normal((symb.n, 2))
    BN = B.copy()
    on.trsm(symb, L.blkval.numpy(), BN, trans="N")
    ip = symb.ip
    ref_N = np.linalg.solve(Ld, B[symb.p])[ip]
    assert np.abs(BN - ref_N).max() < 1e-12 * np.abs(ref_N).max()
    BT = B.copy()
    on.trsm(symb, L.blkval.numpy(), BT, trans="T")
    ref_T = np.linalg.solve(Ld.T, B[symb.p])[ip]
    assert np.abs(BT - ref_T).max() < 1e-12 * np.abs(ref_T).max()
    # products: alpha L b and alpha L^T b in the original ordering, and the
    # product undone by the solve
    for tr, Lop in (("N", Ld), ("T", Ld.T)):
        BM = B.copy()
        on.trmm(symb, L.blkval.numpy(), BM, alpha=1.5, trans=tr)
        ref_M = (1.5 * Lop @ B[symb.p])[ip]
        assert np.abs(BM - ref_M).max() < 1e-12 * np.abs(ref_M).max()
        on.trsm(symb, L.blkval.numpy(), BM, alpha=1.0 / 1.5, trans=tr)
        assert np.abs(BM - B).max() < 1e-12 * np.abs(B).max()
    # llt(cholesky(A)) == A and <P(A^-1), A> = n
    P = L.blkval.numpy().copy()
    on.llt(symb, P)
    assert np.abs(P - b).max() < 1e-12 * np.abs(b).max()
    assert abs(on.dot(symb, Y, b) - symb.n) < 1e-9 * symb.n
    Xd = cp.cspmatrix(symb, blkval=b, device="cpu").spmatrix(reordered=True, symmetric=True).toarray()
    assert abs(on.dot(symb, b, Y) - np.sum(Xd * Ys)) < 1e-9 * symb.n


def test_planted_mistake_is_caught():
    """A copy of the oracle with the Schur-complement update dropped must
    fail the closed-form check (the checks above are not vacuous)."""
    import oracle.numeric as m
    A = problems.grid2d(6, seed=0)
    symb = cp.symbolic(A, p=cp.amd.order)
    b = (cp.cspmatrix(symb, device="cpu") + A).blkval.numpy().copy()
    orig = m._extend_add
    try:
        m._extend_add = lambda F, U, r: None
        bad = b.copy()
        m.cholesky(symb, bad)
    finally:
        m._extend_add = orig
    good = b.copy()
    m.cholesky(symb, good)
    assert _rel(bad, good) > 1e-3


def _factor_and_inverse(symb, A):
    b = (cp.cspmatrix(symb, device="cpu") + A).blkval.numpy().copy()
    L = b.copy()
    assert on.cholesky(symb, L) == 0
    Y = L.copy()
    assert on.projected_inverse(symb, Y) == 0
    return b, L, Y


def _dense_sym(symb, blkval):
    return cp.cspmatrix(symb, blkval=blkval, device="cpu").spmatrix(
        reordered=True, symmetric=True).toarray()


def _lower_mask(symb):
    """True on the meaningful (lower-trapezoid) entries of blkval."""
    m = np.zeros(int(symb.blkptr[-1]), dtype=bool)
    for k in range(symb.Nsn):
        nn = int(symb.snptr[k + 1] - symb.snptr[k])
        nj = nn + int(symb.relptr[k + 1] - symb.relptr[k])
        blk = m[symb.blkptr[k]:symb.blkptr[k + 1]].reshape((nn, nj)).T
        blk[np.tril_indices(nj, 0, nn)] = True
    return m


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("maker", [lambda: problems.reference_test_matrix(),
                                   lambda: problems.grid2d(14, seed=0),
                                   lambda: problems.grid3d(6, seed=1)])
def test_completion_hessian_vs_reference_kernels(maker):
    A = maker()
    symb = cp.symbolic(A, p=cp.amd.order)
    b, L, Y = _factor_and_inverse(symb, A)
    for fu in (True, False):
        x1, x2 = Y.copy(), Y.copy()
        assert ref.completion(symb, x1, fu) == 0
        assert on.completion(symb, x2, fu) == 0
        assert _rel(x2, x1) < 1e-12
    rng = np.random.default_rng(0)
    for fu in (True, False):
        for adj in (False, True, None):
            for inv in (False, True):
                us = [rng.standard_normal(len(b)) for _ in range(2)]
                u1 = [u.copy() for u in us]
                u2 = [u.copy() for u in us]
                ref.hessian(symb, L, Y, u1, adj=adj, inv=inv, factored_updates=fu)
                on.hessian(symb, L, Y, u2, adj=adj, inv=inv, factored_updates=fu)
                for a, c in zip(u2, u1):
                    assert _rel(a, c) < 1e-12, (fu, adj, inv)


def test_update_factor_closed_form():
    """Li^T Li = F[r, r] when the (2,2) block of F holds Lk with
    V_k = Lk^T Lk; Li is lower triangular."""
    rng = np.random.default_rng(3)
    nn, na = 5, 9
    nj = nn + na
    M = rng.standard_normal((nj, nj))
    Vfull = M @ M.T + nj * np.eye(nj)
    Lk = np.linalg.cholesky(Vfull[nn:, nn:][::-1, ::-1]).T[::-1, ::-1]  # V22 = Lk^T Lk
    F = np.tril(Vfull)
    F[nn:, nn:] = np.tril(Lk)
    for r in ([0, 2, 3, 6, 7, 10, 13], [1, 4], [5, 8, 12], [0, 1, 2, 3, 4]):
        r = np.array(r)
        Li = on.update_factor(r, nn, na, F)
        assert np.allclose(np.triu(Li, 1), 0.0)
        assert np.abs(Li.T @ Li - Vfull[np.ix_(r, r)]).max() < 1e-12 * np.abs(Vfull).max()


@pytest.mark.parametrize("maker", [lambda: problems.reference_test_matrix(),
                                   lambda: problems.grid2d(10, seed=0),
                                   lambda: problems.grid3d(5, seed=2)])
def test_completion_hessian_closed_forms(maker):
    """completion(P(X^{-1})) = L (tests/test_numeric.py:48-64 of the
    reference); the full Hessian map is P(X^{-1} U X^{-1}) (dense); G and its
    inverse compose to the identity and G^adj is the adjoint of G under the
    trace inner product (tests/test_numeric.py:66-90)."""
    A = maker()
    symb = cp.symbolic(A, p=cp.amd.order)
    b, L, Y = _factor_and_inverse(symb, A)
    mask = _lower_mask(symb)
    for fu in (True, False):
        Lc = Y.copy()
        assert on.completion(symb, Lc, fu) == 0
        assert np.abs(Lc - L)[mask].max() < 1e-12 * np.abs(L).max()
    PAP = _dense(A)[np.ix_(symb.p, symb.p)]
    Xi = np.linalg.inv(PAP)
    Pat = symb.sparsity_pattern(reordered=True, symmetric=True).toarray() != 0
    rng = np.random.default_rng(5)
    U0 = rng.standard_normal(len(b))
    Ud = _dense_sym(symb, U0)
    for fu in (True, False):
        H = U0.copy()
        on.hessian(symb, L, Y, [H], adj=None, factored_updates=fu)
        Hd = Xi @ Ud @ Xi
        assert np.abs(_dense_sym(symb, H)[Pat] - Hd[Pat]).max() < 1e-11 * np.abs(Hd).max()
        Hi = H.copy()
        on.hessian(symb, L, Y, [Hi], adj=None, inv=True, factored_updates=fu)
        assert np.abs(Hi - U0)[mask].max() < 1e-10 * np.abs(U0).max()
        for adj in (False, True):
            G = U0.copy()
            on.hessian(symb, L, Y, [G], adj=adj, factored_updates=fu)
            on.hessian(symb, L, Y, [G], adj=adj, inv=True, factored_updates=fu)
            assert np.abs(G - U0)[mask].max() < 1e-10 * np.abs(U0).max()
        V0 = rng.standard_normal(len(b))
        GU, GaV = U0.copy(), V0.copy()
        on.hessian(symb, L, Y, [GU], adj=False, factored_updates=fu)
        on.hessian(symb, L, Y, [GaV], adj=True, factored_updates=fu)
        lhs = np.sum(_dense_sym(symb, GU) * _dense_sym(symb, V0))
        rhs = np.sum(_dense_sym(symb, U0) * _dense_sym(symb, GaV))
        assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0)


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
def test_factored_hessian_sign_quirk():
    """With factored updates the reference's G_X depends on the signs of
    Householder pivots that vanish structurally (|alpha| ~ 1e-23 ||x|| on a
    40 x 40 grid), i.e. on BLAS rounding; the oracle takes such pivots as +0.
    G_X is then the reference's up to row signs of the update factors, so
    every sign-invariant quantity agrees: <G(U), G(V)>, G^adj G and G^{-1} G."""
    A = problems.grid2d(40, seed=0)
    symb = cp.symbolic(A, p=cp.amd.order)
    b, L, Y = _factor_and_inverse(symb, A)
    rng = np.random.default_rng(2)
    U0, V0 = rng.standard_normal(len(b)), rng.standard_normal(len(b))
    gr = [U0.copy(), V0.copy()]
    go = [U0.copy(), V0.copy()]
    ref.hessian(symb, L, Y, gr, adj=False, factored_updates=True)
    on.hessian(symb, L, Y, go, adj=False, factored_updates=True)
    d = lambda x, y: on.dot(symb, x, y)
    assert abs(d(*gr) - d(*go)) < 1e-11 * abs(d(*gr))
    assert abs(d(gr[0], gr[0]) - d(go[0], go[0])) < 1e-11 * d(gr[0], gr[0])
    hr, ho = [U0.copy()], [U0.copy()]
    ref.hessian(symb, L, Y, hr, adj=None, factored_updates=True)
    on.hessian(symb, L, Y, ho, adj=None, factored_updates=True)
    assert _rel(ho[0], hr[0]) < 1e-12


def test_pf_oracle_closed_form():
    """The product-form factors reproduce I + V diag(a) V^T: the dense
    product (L_1 ... L_m)(L_1 ... L_m)^T with L_i = diag(l_i) +
    strictly-lower(v_i b_i^T), and the solve / product routines invert each
    other and agree with the dense factor."""
    rng = np.random.default_rng(4)
    n, m = 40, 5
    V0 = rng.standard_normal((n, m))
    a = rng.uniform(0.5, 2.0, m)
    a[2] = -0.01   # a small downdate stays positive definite
    V, L, B = V0.copy(), np.zeros((n, m)), np.zeros((n, m))
    assert on.pf_factor(a, V, L, B) == 0
    F = np.eye(n)
    for i in range(m):
        Li = np.diag(L[:, i]) + np.tril(np.outer(V[:, i], B[:, i]), -1)
        F = F @ Li
    M = np.eye(n) + V0 @ np.diag(a) @ V0.T
    assert np.abs(F @ F.T - M).max() < 1e-12 * np.abs(M).max()
    U0 = rng.standard_normal((n, 3))
    for tr, op in (("N", F), ("T", F.T)):
        U = U0.copy()
        on.pf_product(V, L, B, U, trans=tr)
        assert np.abs(U - op @ U0).max() < 1e-12 * np.abs(op @ U0).max()
        on.pf_solve(V, L, B, U, trans=tr)
        assert np.abs(U - U0).max() < 1e-12 * np.abs(U0).max()
    # a downdate that makes I + V diag(a) V^T indefinite fails
    V2 = V0.copy()
    assert on.pf_factor(np.array([-10.0] * m), V2, np.zeros((n, m)), np.zeros((n, m))) == -1


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
def test_pf_oracle_vs_reference_kernels():
    """The oracle's product-form routines against pfcholesky.c itself
    (oracle/_ref): the same sequential arithmetic, so equal to the last bit
    up to the compiler's contraction of a*b+c."""
    rng = np.random.default_rng(9)
    n, m = 300, 4
    V0 = np.asfortranarray(rng.standard_normal((n, m)))
    a = np.array([1.0, 0.5, -0.002, 2.0])
    V1, L1, B1 = V0.copy(order="F"), np.zeros((n, m), order="F"), np.zeros((n, m), order="F")
    V2, L2, B2 = V0.copy(order="F"), np.zeros((n, m), order="F"), np.zeros((n, m), order="F")
    assert on.pf_factor(a, V1, L1, B1) == 0
    assert ref.pf_factor(a, V2, L2, B2) == 0
    for x, y in ((V1, V2), (L1, L2), (B1, B2)):
        assert _rel(x, y) < 1e-14
    U0 = np.asfortranarray(rng.standard_normal((n, 3)))
    for tr in "NT":
        for ofun, rfun in ((on.pf_solve, ref.pf_solve), (on.pf_product, ref.pf_product)):
            Uo, Ur = U0.copy(order="F"), U0.copy(order="F")
            ofun(V1, L1, B1, Uo, trans=tr)
            rfun(V2, L2, B2, Ur, trans=tr)
            assert _rel(Uo, Ur) < 1e-14
    assert ref.pf_factor(np.full(m, -10.0), V0.copy(order="F"), np.zeros((n, m), order="F"),
                         np.zeros((n, m), order="F")) == -1
