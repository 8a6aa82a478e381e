"""Chordal orderings: maximum cardinality search and the maximal chordal
subgraph (C++ in csrc/symbolic.cpp) against the plain-Python restatement of
mcs.py / maxchord.py (exact), and both against what the mathematics fixes:
the reference documentation's 17 x 17 example (quick_start.rst:255-297: a
perfect elimination ordering without fill), chordality and brute-force
maximality of the subgraph, maxchord(A) = A for chordal A."""
import numpy as np
import pytest
import scipy.sparse as sp

import chompack_amd as cp
from oracle import orderings as oo


def _doc17():
    I = list(range(17)) + [2, 2, 3, 3, 4, 14, 4, 14, 8, 14, 15, 8, 15, 7, 8, 14, 8, 14, 14,
                           15, 10, 12, 13, 16, 12, 13, 16, 12, 13, 15, 16, 13, 15, 16, 15, 16,
                           15, 16, 16]
    J = list(range(17)) + [0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 9, 9,
                           9, 9, 10, 10, 10, 11, 11, 11, 11, 12, 12, 12, 13, 13, 14, 14, 15]
    return sp.csc_matrix((np.ones(len(I)), (I, J)), shape=(17, 17))


def _random_graph(n, dens, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=dens, random_state=seed, format="coo")
    I, J = M.row, M.col
    keep = I > J
    I = np.concatenate([I[keep], np.arange(n)])
    J = np.concatenate([J[keep], np.arange(n)])
    return sp.csc_matrix((rng.uniform(0.5, 1.5, len(I)), (I, J)), shape=(n, n))


def _chordal(n, dens, seed):
    """The filled pattern of a random matrix (chordal), values of A where
    present and 1 on fill."""
    A = _random_graph(n, dens, seed)
    symb = cp.symbolic(A, p=cp.amd.order)
    S = symb.sparsity_pattern(reordered=False, symmetric=False)
    return sp.csc_matrix(sp.tril(S))


def _is_chordal(A):
    return cp.peo(A, cp.maxcardsearch(A))


def test_doc_example_17():
    A = _doc17()
    p = cp.maxcardsearch(A)
    assert cp.peo(A, p)
    assert tuple(int(x) for x in cp.symbolic(A, p=p).fill) == (0, 0)
    assert np.array_equal(p, oo.maxcardsearch(A))


@pytest.mark.parametrize("seed", range(6))
def test_maxcardsearch_matches_restatement(seed):
    for A in (_random_graph(60, 0.05, seed), _chordal(80, 0.03, seed)):
        n = A.shape[0]
        for ve in (None, 0, n // 2):
            assert np.array_equal(cp.maxcardsearch(A, ve), oo.maxcardsearch(A, ve))


@pytest.mark.parametrize("seed", range(6))
def test_maxcardsearch_is_peo_of_chordal(seed):
    A = _chordal(120, 0.02, seed)
    p = cp.maxcardsearch(A)
    assert sorted(p.tolist()) == list(range(A.shape[0]))
    assert cp.peo(A, p)
    assert tuple(int(x) for x in cp.symbolic(A, p=p).fill) == (0, 0)


@pytest.mark.parametrize("seed", range(6))
def test_maxchord(seed):
    A = _random_graph(30, 0.15, seed)
    Ac, p = cp.maxchord(A)
    Ar, pr = oo.maxchord(A)
    assert np.array_equal(p, pr)
    assert (Ac != Ar).nnz == 0 and np.array_equal(Ac.toarray(), Ar.toarray())
    # a subgraph of A, with A's values, chordal with p as elimination order
    D = A.toarray()
    Cd = Ac.toarray()
    nzc = Cd != 0
    assert np.all(np.tril(nzc) == nzc)
    assert np.array_equal(Cd[nzc], D[nzc])
    assert cp.peo(Ac, p)
    # maximal: adding any dropped edge of A breaks chordality
    dropped = np.argwhere(np.tril(D != 0, -1) & ~nzc)
    for i, j in dropped:
        E = Cd.copy()
        E[i, j] = 1.0
        assert not _is_chordal(sp.csc_matrix(E)), (i, j)


def test_maxchord_of_chordal_is_identity():
    for seed in range(3):
        A = _chordal(100, 0.03, seed)
        Ac, p = cp.maxchord(A)
        assert np.array_equal(Ac.toarray(), A.toarray())
        assert cp.peo(A, p)
