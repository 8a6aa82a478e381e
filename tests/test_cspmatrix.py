"""cspmatrix assembly / export round trips (tests/test_cspmatrix.py:16-46 of
the reference), on host (CPU) blkval arrays."""
import numpy as np

import chompack_amd as cp
from tests.test_symbolic import A17


def test_cspmatrix_roundtrips():
    A = A17()
    symb = cp.symbolic(A, p=cp.amd.order)
    Ac = cp.cspmatrix(symb, device="cpu") + A
    assert Ac.is_factor is False
    assert Ac.size == (17, 17)
    d = (A - Ac.spmatrix(reordered=False, symmetric=False)).toarray()
    assert np.abs(d).max() == 0
    d = (cp.symmetrize(A) - Ac.spmatrix(reordered=False, symmetric=True)).toarray()
    assert np.abs(d).max() == 0
    S = cp.symmetrize(A).toarray()[np.ix_(symb.p, symb.p)]
    assert np.abs(S - Ac.spmatrix(reordered=True, symmetric=True).toarray()).max() == 0
    assert np.abs(np.tril(S) - Ac.spmatrix(reordered=True, symmetric=False).toarray()).max() == 0
    assert np.abs((2 * A).toarray() - (2 * Ac).spmatrix(reordered=False).toarray()).max() == 0
    assert np.abs((2 * A).toarray() - (Ac + Ac).spmatrix(reordered=False).toarray()).max() == 0
    Ac2 = Ac.copy()
    Ac2 += Ac
    assert np.abs((2 * A).toarray() - Ac2.spmatrix(reordered=False).toarray()).max() == 0
    assert np.array_equal(Ac.diag(reordered=False), A.diagonal())
    assert np.array_equal(Ac.diag(reordered=True), A.diagonal()[symb.p])


def test_add_projection_and_pattern():
    A = A17()
    symb = cp.symbolic(A, p=cp.amd.order)
    Xd = np.random.default_rng(0).standard_normal((17, 17))
    Xd = Xd + Xd.T
    X = cp.cspmatrix(symb, device="cpu")
    X.add_projection(Xd)
    P = symb.sparsity_pattern(reordered=False, symmetric=True).toarray() != 0
    E = X.spmatrix(reordered=False, symmetric=True).toarray()
    assert np.allclose(E[P], Xd[P])
    assert np.all(E[~P] == 0)
