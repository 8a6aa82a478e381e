"""Symbolic layer (CPU, integer): reference test goldens, documentation
goldens, and brute-force dense elimination.

Goldens: tests/test_symbolic.py:9-98 of the reference (17x17 chordal pattern,
23x23 non-chordal A_nc with natural and AMD orderings, merge heuristics) and
doc/source/quick_start.rst:40-127,178-184 (7x7 example: permutation, cliques,
clique tree, supernodes, separators, reordered L)."""
import numpy as np
import pytest
import scipy.sparse as sp

import chompack_amd as cp
from chompack_amd import problems

I17 = list(range(17)) + [2, 2, 3, 3, 4, 14, 4, 14, 8, 14, 15, 8, 15, 7, 8, 14, 8, 14, 14, 15, 10,
                         12, 13, 16, 12, 13, 16, 12, 13, 15, 16, 13, 15, 16, 15, 16, 15, 16, 16]
J17 = list(range(17)) + [0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 9, 9, 9, 9,
                         10, 10, 10, 11, 11, 11, 11, 12, 12, 12, 13, 13, 14, 14, 15]
NZ23 = [1, 2, 10, 12, 13, 14, 15, 16, 19, 24, 48, 58, 71, 76, 81, 85, 91, 103, 108, 109, 111,
        114, 116, 117, 118, 134, 143, 145, 161, 174, 178, 180, 183, 192, 194, 202, 203, 205,
        212, 214, 219, 224, 226, 227, 228, 229, 235, 240, 241, 243, 247, 255, 256, 260, 269,
        273, 275, 279, 280, 314, 315, 320, 321, 328, 340, 342, 344, 345, 349, 350, 357, 359,
        370, 372, 375, 384, 385, 394, 399, 402, 411, 412, 419, 420, 422, 433, 439, 441, 447,
        452, 454, 458, 460, 474, 476, 479, 481, 483, 485, 497, 506, 517, 519, 523, 526, 527]


def A17():
    return sp.csc_matrix((np.arange(len(I17), dtype=float), (I17, J17)), shape=(17, 17))


def A23():
    n = 23
    v = np.arange(1, len(NZ23) + 1) / len(NZ23)
    A = sp.csc_matrix((v, ([x % n for x in NZ23], [x // n for x in NZ23])), shape=(n, n))
    return cp.tril(A) + sp.identity(n, format="csc") * 10.0


def test_peo():
    assert cp.peo(A17(), np.arange(17))


def test_symbolic_17():
    s = cp.symbolic(A17(), p=None)
    assert (s.n, s.nnz, s.clique_number, s.Nsn, s.fill) == (17, 56, 5, 9, (0, 0))


def test_merge_17():
    s = cp.symbolic(A17(), p=None, merge_function=cp.merge_size_fill(0, 0))
    assert (s.n, s.nnz, s.Nsn, s.fill, s.clique_number) == (17, 56, 9, (0, 0), 5)
    s = cp.symbolic(A17(), p=None, merge_function=cp.merge_size_fill(2, 2))
    assert s.n == 17 and s.nnz > 56 and s.Nsn < 9 and s.fill[0] >= 0 and s.fill[1] > 0


def test_symbolic_nc():
    s = cp.symbolic(A23(), p=None)
    assert (s.n, s.nnz, s.clique_number, s.Nsn, s.fill) == (23, 150, 12, 10, (73, 0))
    s = cp.symbolic(A23(), p=cp.amd.order)
    assert (s.n, s.nnz, s.clique_number, s.Nsn, s.fill) == (23, 113, 9, 15, (36, 0))


def test_merge_nc():
    s = cp.symbolic(A23(), p=None, merge_function=cp.merge_size_fill(0, 0))
    assert (s.n, s.nnz, s.clique_number, s.Nsn, s.fill) == (23, 150, 12, 10, (73, 0))
    s = cp.symbolic(A23(), p=cp.amd.order, merge_function=cp.merge_size_fill(4, 4))
    assert s.n == 23 and s.nnz > 150 and s.Nsn < 10 and s.fill[0] >= 36 and s.fill[1] > 0


def test_quick_start_7x7():
    I = [0, 1, 3, 1, 5, 2, 6, 3, 4, 5, 4, 5, 6, 5, 6]
    J = [0, 0, 0, 1, 1, 2, 2, 3, 3, 3, 4, 4, 4, 5, 6]
    A = sp.csc_matrix((np.arange(1.0, 16.0), (I, J)), shape=(7, 7))
    s = cp.symbolic(A, p=cp.amd.order)
    assert list(s.p) == [2, 6, 4, 1, 3, 5, 0]
    assert s.cliques() == [[0, 1], [1, 2], [2, 4, 5], [3, 5, 6], [4, 5, 6]]
    assert s.parent() == [1, 2, 4, 4, 4]
    assert s.supernodes() == [[0], [1], [2], [3], [4, 5, 6]]
    assert s.separators() == [[1], [2], [4, 5], [5, 6], []]
    L = (cp.cspmatrix(s, device="cpu") + A).spmatrix(reordered=True).toarray()
    expect = np.array([[6, 0, 0, 0, 0, 0, 0], [7, 15, 0, 0, 0, 0, 0], [0, 13, 11, 0, 0, 0, 0],
                       [0, 0, 0, 4, 0, 0, 0], [0, 0, 9, 0, 8, 0, 0], [0, 0, 12, 5, 10, 14, 0],
                       [0, 0, 0, 2, 3, 0, 1]], dtype=float)
    assert np.array_equal(L, expect)


def _brute_force_fill(A, p):
    """Dense symbolic elimination of P A P^T: boolean lower pattern of L."""
    S = (cp.symmetrize(A).toarray() != 0)
    S = S[np.ix_(p, p)]
    n = S.shape[0]
    for k in range(n):
        nz = np.nonzero(S[k + 1:, k])[0] + k + 1
        if len(nz):
            S[np.ix_(nz, nz)] = True
    return np.tril(S)


@pytest.mark.parametrize("seed", range(6))
def test_symbolic_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(8, 60))
    m = int(rng.integers(n, 4 * n))
    I = rng.integers(0, n, m)
    J = rng.integers(0, n, m)
    A = sp.csc_matrix((np.ones(m), (np.maximum(I, J), np.minimum(I, J))), shape=(n, n)) \
        + sp.identity(n, format="csc")
    for p in (None, cp.amd.order):
        s = cp.symbolic(A, p=p)
        F = _brute_force_fill(A, s.p)
        # clique k's rows are exactly the pattern of the columns of supernode k
        for k in range(s.Nsn):
            rows = s.snrowidx[s.sncolptr[k]:s.sncolptr[k + 1]]
            for c in range(s.snptr[k], s.snptr[k + 1]):
                col = np.nonzero(F[:, c])[0]
                assert list(col) == [r for r in rows if r >= c]
        assert s.nnz == int(F.sum())
        # relative indices address the parent's clique
        for k in range(s.Nsn):
            pk = s.snpar[k]
            nn = s.snptr[k + 1] - s.snptr[k]
            sep = s.snrowidx[s.sncolptr[k] + nn:s.sncolptr[k + 1]]
            rel = s.relidx[s.relptr[k]:s.relptr[k + 1]]
            if pk == k:
                assert len(sep) == 0
            else:
                pr = s.snrowidx[s.sncolptr[pk]:s.sncolptr[pk + 1]]
                assert list(pr[rel]) == list(sep)
        # postorder: children before parents, supernodes contiguous
        assert all(s.snpar[k] >= k for k in range(s.Nsn))
        assert np.array_equal(np.sort(s.p), np.arange(n))


def test_supernodes_are_maximal_fundamental():
    """Pothen-Sun rule: j's etree parent joins j's supernode iff
    colcount[j] - 1 == colcount[parent] and no earlier sibling claimed it."""
    A = problems.grid2d(10, seed=0)
    s = cp.symbolic(A, p=cp.amd.order)
    Ap = cp.perm(cp.symmetrize(A), s.p)
    par = cp.symbolic.__module__ and __import__("chompack_amd.symbolic", fromlist=["etree"]).etree(Ap)
    from chompack_amd.symbolic import counts, post_order
    post = post_order(par)
    cc = counts(Ap, par, post)
    F = _brute_force_fill(A, s.p)
    assert np.array_equal(cc, F.sum(axis=0))
    for k in range(s.Nsn):
        cols = range(s.snptr[k], s.snptr[k + 1])
        for c in list(cols)[:-1]:
            assert par[c] == c + 1 and cc[c] - 1 == cc[c + 1]


def test_memory_plan_matches_golden():
    from pathlib import Path
    g = np.load(Path(__file__).parent / "golden" / "ref31_amd.npz")
    s = cp.symbolic(problems.reference_test_matrix(), p=cp.amd.order)
    for key in ("p", "snptr", "snpar", "relptr", "relidx", "chptr", "chidx", "blkptr",
                "sncolptr", "snrowidx"):
        assert np.array_equal(np.asarray(getattr(s, key)), g["symb_" + key]), key
    mem = [s.memory[k] for k in ("stack_depth", "stack_mem", "frontal_mem", "stack_solve")]
    assert mem == list(g["symb_memory"])


def test_amd_fill_quality():
    """AMD must beat the natural order by a wide margin on a 3D grid and be a
    valid permutation; an MMD ordering (SuperLU) is the yardstick."""
    import scipy.sparse.linalg as sla
    A = problems.grid3d(10, seed=1)
    p = cp.amd.order(A)
    assert np.array_equal(np.sort(p), np.arange(A.shape[0]))
    s_amd = cp.symbolic(A, p=p)
    s_nat = cp.symbolic(A, p=None)
    S = cp.symmetrize(A).tocsc()
    lu = sla.splu(S, permc_spec="MMD_AT_PLUS_A", options={"SymmetricMode": True},
                  diag_pivot_thresh=0.0)
    s_mmd = cp.symbolic(A, p=lu.perm_c)
    assert s_amd.nnz < 0.6 * s_nat.nnz
    assert s_amd.nnz < 1.25 * s_mmd.nnz


def _lifo_memory(s):
    """The reference's memory plan by simulating its sequential sweep
    (symbolic.py:606-680): supernodes in order, each popping its children's
    update matrices and pushing its own."""
    nnk = np.diff(s.snptr)
    na = np.diff(s.relptr)
    nch = np.diff(s.chptr)
    stack, mem, solve, depth, cur, cur_s = [], 0, 0, 0, 0, 0
    for k in range(s.Nsn):
        for _ in range(int(nch[k])):
            v = stack.pop()
            cur -= v * v
            cur_s -= v
        if na[k] > 0:
            stack.append(int(na[k]))
            cur += int(na[k]) ** 2
            cur_s += int(na[k])
            mem, solve = max(mem, cur), max(solve, cur_s)
        depth = max(depth, len(stack))
    cln = int((nnk + na).max())
    return {"stack_depth": depth, "stack_mem": mem, "frontal_mem": cln * cln, "stack_solve": solve}


@pytest.mark.parametrize("make,order", [
    (lambda: problems.grid3d(14, seed=1), "amd"), (lambda: problems.grid2d(40, seed=0), "amd"),
    (lambda: problems.caterpillar(), None), (lambda: problems.two_cliques(), None),
    (lambda: A23(), "amd"), (lambda: A17(), None)])
def test_memory_plan_matches_lifo_simulation(make, order):
    s = cp.symbolic(make(), p=cp.amd.order if order else None)
    assert s.memory == _lifo_memory(s)
