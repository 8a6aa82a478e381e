"""Dense completions: psdcompletion and mrcompletion.

CPU tests pin the oracle (oracle/completions.py) by what the mathematics
fixes: the maximum-determinant completion X reproduces A on the pattern and
its inverse vanishes off the pattern, and psdcompletion(P(A^{-1})) = A^{-1};
with a zero row and column (the pseudo-inverse branch) the rest is the
maximum-determinant completion of the reduced matrix; the minimum-rank
completion reproduces A through Y Y^T with r = the largest clique rank (the
reference test's own check, tests/test_mrcompletion.py:18-36 of the
reference, on its 17 x 17 example); the EDM completion of squared distances
reproduces them on the pattern and is a Euclidean distance matrix.

GPU tests compare the HIP path with the oracle: psdcompletion entrywise,
through the multifrontal route and through the reference's sweep
(max |gpu - oracle| / max |oracle| <= 1e-12 on the Cholesky branch, 1e-11
on the pseudo-inverse branch, where the eigensolvers differ), edmcompletion
entrywise (1e-11, pseudo-inverses throughout), mrcompletion
through P(Y Y^T) = A (1e-12 relative) and the rank (Y itself is only
determined up to an orthogonal factor).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import chompack_amd as cp
from chompack_amd import problems
from oracle import completions as oc
from tests.test_symbolic import I17, J17


def _blkval(symb, A):
    X = cp.cspmatrix(symb, device="cpu") + A
    return np.asarray(X.blkval, dtype=np.float64).copy()


def _pattern_re(symb):
    """Boolean n x n mask of the filled pattern (reordered, both triangles)."""
    return oc.dense_projection(symb, np.ones(int(symb.blkptr[-1]))) != 0


def _proj_lower(symb, M):
    """Lower-triangular sparse matrix with M's entries (original numbering)
    on the filled pattern."""
    P = sp.coo_matrix(symb.sparsity_pattern(reordered=False, symmetric=False))
    return sp.csc_matrix((M[P.row, P.col], (P.row, P.col)), shape=M.shape)


def _ref17():
    A = sp.csc_matrix((np.ones(len(I17)), (I17, J17)), shape=(17, 17)).tolil()
    A[0, 0] += 1.0
    A[1, 1] += 1.0
    return sp.csc_matrix(A)


def _offpattern_inverse(X, mask):
    Xi = np.linalg.inv(X)
    return np.abs(Xi[~mask]).max(initial=0.0) / np.abs(Xi).max()


def _clique_rank(symb, blk):
    Xd = oc.dense_projection(symb, blk)
    r = 0
    for k in range(symb.Nsn):
        w = np.linalg.eigvalsh(oc.clique_matrix(symb, blk, k, Xd))
        r = max(r, int(np.sum(w > 1e-14 * w[-1])))
    return r


# ---------------------------------------------------------------- oracle pins

@pytest.mark.parametrize("side", [6, 9])
def test_psd_oracle_maxdet(side):
    A = problems.grid2d(side, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    blk = _blkval(symb, A)
    X = oc.psdcompletion(symb, blk)
    mask = _pattern_re(symb)
    Ad = oc.dense_projection(symb, blk)
    assert np.abs(X - X.T).max() == 0.0
    assert np.abs((X - Ad)[mask]).max() <= 1e-14 * np.abs(Ad).max()
    assert _offpattern_inverse(X, mask) <= 1e-11
    assert np.linalg.eigvalsh(X).min() > 0
    # reordered=False: X[ip, ip], which reproduces A in the original numbering
    Xo = oc.psdcompletion(symb, blk, reordered=False)
    Ao = (cp.symmetrize(A)).toarray()
    nz = Ao != 0
    assert np.abs((Xo - Ao)[nz]).max() <= 1e-14 * np.abs(Ao).max()


@pytest.mark.parametrize("side", [5, 8])
def test_psd_oracle_closed_form(side):
    """The inverse of an SPD matrix with the chordal pattern is the
    maximum-determinant completion of its own projection: psdcompletion
    (P(A^{-1})) = A^{-1}, dense."""
    A = problems.grid2d(side, seed=6)
    symb = cp.symbolic(A, p=cp.amd.order)
    Ad = oc.dense_projection(symb, _blkval(symb, A))
    Ainv = np.linalg.inv(Ad)
    blk = _blkval(symb, _proj_lower(symb, np.linalg.inv(cp.symmetrize(A).toarray())))
    X = oc.psdcompletion(symb, blk)
    assert np.abs(X - Ainv).max() <= 1e-13 * np.abs(Ainv).max()
    assert (X != 0).all()


def test_psd_oracle_pseudo_inverse_branch():
    """A zero row and column in the root supernode makes every separator that
    holds it singular (the reference's eigendecomposition branch): the row
    stays zero and the rest is the max-det completion of the reduced
    matrix."""
    A = problems.grid2d(7, seed=4)
    symb = cp.symbolic(A, p=cp.amd.order)
    v_re = symb.n - 1
    v = int(np.asarray(symb.p)[v_re])
    A0 = A.tolil()
    A0[v, :] = 0.0
    A0[:, v] = 0.0
    A0 = sp.csc_matrix(A0)
    blk = _blkval(symb, A0)
    X = oc.psdcompletion(symb, blk)
    assert np.abs(X[v_re]).max() == 0.0 and np.abs(X[:, v_re]).max() == 0.0
    keep = np.arange(symb.n) != v_re
    Xr = X[np.ix_(keep, keep)]
    mask = _pattern_re(symb)[np.ix_(keep, keep)]
    Ad = oc.dense_projection(symb, blk)[np.ix_(keep, keep)]
    assert np.abs((Xr - Ad)[mask]).max() <= 1e-14 * np.abs(Ad).max()
    assert _offpattern_inverse(Xr, mask) <= 1e-11


def test_mr_oracle_reference_example():
    A = _ref17()
    symb = cp.symbolic(A, p=cp.amd.order)
    blk = _blkval(symb, A)
    mask = _pattern_re(symb)
    Ad = oc.dense_projection(symb, blk)
    Y = oc.mrcompletion(symb, blk)
    assert Y.shape[1] == _clique_rank(symb, blk)
    assert np.abs((Y @ Y.T - Ad)[mask]).max() <= 1e-12
    # the rank-2 case of the reference test: A = P(U U^T)
    U = np.random.default_rng(17).standard_normal((17, 2))
    A2 = _proj_lower(symb, U @ U.T)
    blk2 = _blkval(symb, A2)
    Y2 = oc.mrcompletion(symb, blk2)
    assert Y2.shape[1] == 2
    Ad2 = oc.dense_projection(symb, blk2)
    assert np.abs((Y2 @ Y2.T - Ad2)[mask]).max() <= 1e-12 * np.abs(Ad2).max()
    Y2o = oc.mrcompletion(symb, blk2, reordered=False)
    Po = symb.sparsity_pattern(reordered=False, symmetric=True).toarray() != 0
    assert np.abs((Y2o @ Y2o.T - U @ U.T)[Po]).max() <= 1e-12 * np.abs(U @ U.T).max()


def _edm_problem(side, seed):
    """Squared distances of random points on the filled pattern of a 2D
    grid; the dimension exceeds every separator, so each Gram matrix of a
    separator has exactly one zero eigenvalue (its first row)."""
    A = problems.grid2d(side, seed=seed)
    symb = cp.symbolic(A, p=cp.amd.order)
    d = int(max(np.diff(symb.sncolptr) - np.diff(symb.snptr))) + 4
    pts = np.random.default_rng(seed).standard_normal((symb.n, d))
    D = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    return symb, _proj_lower(symb, D)


def _is_edm(X):
    n = X.shape[0]
    J = np.eye(n) - np.ones((n, n)) / n
    G = -0.5 * J @ X @ J
    return np.abs(np.diag(X)).max() == 0.0 and np.linalg.eigvalsh(G).min() >= -1e-10 * np.abs(G).max()


@pytest.mark.parametrize("side", [6, 9])
def test_edm_oracle(side):
    symb, A = _edm_problem(side, side)
    blk = _blkval(symb, A)
    X = oc.edmcompletion(symb, blk)
    mask = _pattern_re(symb)
    Ad = oc.dense_projection(symb, blk)
    assert np.abs((X - Ad)[mask]).max() <= 1e-14 * np.abs(Ad).max()
    assert np.abs(X - X.T).max() == 0.0
    assert _is_edm(X)
    assert (X[~mask] > 0).all()


def test_completions_reject_factor():
    A = problems.grid2d(4, seed=0)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cpu") + A
    X.is_factor = True
    with pytest.raises(AssertionError, match="A must be a cspmatrix"):
        cp.psdcompletion(X)
    with pytest.raises(AssertionError, match="A must be a cspmatrix"):
        cp.mrcompletion(X)
    with pytest.raises(AssertionError, match="A must be a cspmatrix"):
        cp.edmcompletion(X)


# ---------------------------------------------------------------- GPU parity

def _gpu(symb, A):
    import torch
    X = cp.cspmatrix(symb, device="cuda:0") + A
    torch.cuda.synchronize()
    return X


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


PSD_CASES = {
    "grid2d_8": lambda: problems.grid2d(8, seed=0),
    "grid2d_24": lambda: problems.grid2d(24, seed=1),
    "grid3d_6": lambda: problems.grid3d(6, seed=2),
    "ref31": lambda: problems.reference_test_matrix(),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(PSD_CASES))
def test_psd_gpu(name):
    print("case psdcompletion", name, flush=True)
    A = PSD_CASES[name]()
    symb = cp.symbolic(A, p=cp.amd.order)
    X = _gpu(symb, A)
    blk = X.blkval.cpu().numpy()
    ref = oc.psdcompletion(symb, blk)
    got = cp.psdcompletion(X).cpu().numpy()
    e = _rel(got, ref)
    print("psdcompletion rel err %.3e" % e, flush=True)
    assert e <= 1e-12
    got_o = cp.psdcompletion(X, reordered=False).cpu().numpy()
    ip = np.asarray(symb.ip)
    assert _rel(got_o, ref[np.ix_(ip, ip)]) <= 1e-12
    # the reference's sweep (separator Cholesky branch) on the same input
    e = _rel(cp.psdcompletion(X, sweep=True).cpu().numpy(), ref)
    print("psdcompletion sweep rel err %.3e" % e, flush=True)
    assert e <= 1e-12
    assert _rel(cp.psdcompletion(X, reordered=False, sweep=True).cpu().numpy(),
                ref[np.ix_(ip, ip)]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("side", [10, 30])
def test_psd_gpu_closed_form(side):
    """psdcompletion(P(A^{-1})) = A^{-1} (dense) against numpy's inverse
    and the oracle."""
    print("case psdcompletion closed form", side, flush=True)
    A = problems.grid2d(side, seed=6)
    symb = cp.symbolic(A, p=cp.amd.order)
    Ainv_o = np.linalg.inv(cp.symmetrize(A).toarray())
    X = _gpu(symb, _proj_lower(symb, Ainv_o))
    got = cp.psdcompletion(X, reordered=False).cpu().numpy()
    e = _rel(got, Ainv_o)
    print("psdcompletion vs inv(A) rel err %.3e" % e, flush=True)
    assert e <= 1e-12
    ref = oc.psdcompletion(symb, X.blkval.cpu().numpy())
    assert _rel(cp.psdcompletion(X).cpu().numpy(), ref) <= 1e-12


@pytest.mark.gpu
def test_psd_gpu_pseudo_inverse_branch():
    print("case psdcompletion zero vertex", flush=True)
    A = problems.grid2d(10, seed=4)
    symb = cp.symbolic(A, p=cp.amd.order)
    v = int(np.asarray(symb.p)[symb.n - 1])
    A0 = A.tolil()
    A0[v, :] = 0.0
    A0[:, v] = 0.0
    A0 = sp.csc_matrix(A0)
    X = _gpu(symb, A0)
    ref = oc.psdcompletion(symb, X.blkval.cpu().numpy())
    got = cp.psdcompletion(X).cpu().numpy()
    e = _rel(got, ref)
    print("psdcompletion (pseudo-inverse) rel err %.3e" % e, flush=True)
    assert e <= 1e-11
    assert np.abs(got[symb.n - 1]).max() == 0.0


@pytest.mark.gpu
def test_mr_gpu_reference_example():
    print("case mrcompletion 17x17", flush=True)
    A = _ref17()
    symb = cp.symbolic(A, p=cp.amd.order)
    mask = _pattern_re(symb)
    X = _gpu(symb, A)
    blk = X.blkval.cpu().numpy()
    Ad = oc.dense_projection(symb, blk)
    Y = cp.mrcompletion(X).cpu().numpy()
    assert Y.shape[1] == oc.mrcompletion(symb, blk).shape[1]
    assert np.abs((Y @ Y.T - Ad)[mask]).max() <= 1e-12 * np.abs(Ad).max()
    U = np.random.default_rng(17).standard_normal((17, 2))
    X2 = _gpu(symb, _proj_lower(symb, U @ U.T))
    Y2 = cp.mrcompletion(X2).cpu().numpy()
    assert Y2.shape[1] == 2
    Ad2 = oc.dense_projection(symb, X2.blkval.cpu().numpy())
    assert np.abs((Y2 @ Y2.T - Ad2)[mask]).max() <= 1e-12 * np.abs(Ad2).max()
    Y2o = cp.mrcompletion(X2, reordered=False).cpu().numpy()
    Po = symb.sparsity_pattern(reordered=False, symmetric=True).toarray() != 0
    assert np.abs((Y2o @ Y2o.T - U @ U.T)[Po]).max() <= 1e-12 * np.abs(U @ U.T).max()


@pytest.mark.gpu
@pytest.mark.parametrize("side,rank", [(12, 3), (20, 5)])
def test_mr_gpu_low_rank_grid(side, rank):
    print("case mrcompletion grid", side, rank, flush=True)
    A = problems.grid2d(side, seed=5)
    symb = cp.symbolic(A, p=cp.amd.order)
    n = symb.n
    U = np.random.default_rng(side).standard_normal((n, rank))
    X = _gpu(symb, _proj_lower(symb, U @ U.T))
    blk = X.blkval.cpu().numpy()
    Ad = oc.dense_projection(symb, blk)
    mask = _pattern_re(symb)
    Y = cp.mrcompletion(X).cpu().numpy()
    assert Y.shape[1] == rank == oc.mrcompletion(symb, blk).shape[1]
    e = np.abs((Y @ Y.T - Ad)[mask]).max() / np.abs(Ad).max()
    print("mrcompletion P(YY^T) - A rel %.3e" % e, flush=True)
    assert e <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("side", [8, 16])
def test_edm_gpu(side):
    print("case edmcompletion grid", side, flush=True)
    symb, A = _edm_problem(side, side + 1)
    X = _gpu(symb, A)
    blk = X.blkval.cpu().numpy()
    ref = oc.edmcompletion(symb, blk)
    got = cp.edmcompletion(X).cpu().numpy()
    e = _rel(got, ref)
    print("edmcompletion rel err %.3e" % e, flush=True)
    assert e <= 1e-11
    assert _is_edm(got)
    ip = np.asarray(symb.ip)
    assert _rel(cp.edmcompletion(X, reordered=False).cpu().numpy(), ref[np.ix_(ip, ip)]) <= 1e-11
