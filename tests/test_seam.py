"""Plugin-seam helpers on the CPU: ``lmerge`` (cbase.c:64-101,
misc.py:178-204), cvxopt-style sparse and dense arguments (symbolic.py:529
reads a cvxopt ``spmatrix``; cbase.c:1017-1018 a cvxopt ``matrix``), and the
column-major addressing of strided tensors used by trsm / trmm / pfcholesky.
"""
import array

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import chompack_amd as cp
from chompack_amd import misc, numeric, problems
from tests.test_symbolic import A17, A23


# ---------------------------------------------------------------- lmerge
def _multiset_max_union(a, b):
    """For ascending inputs the two-pointer merge keeps each value v
    max(count_a(v), count_b(v)) times: equal heads are taken once per pair,
    the longer run's surplus follows."""
    vals = sorted(set(a) | set(b))
    out = []
    for v in vals:
        out += [v] * max(a.count(v), b.count(v))
    return out


def test_lmerge_cases():
    left = [1, 3, 5, 0, 0, 0, 0]
    k = cp.lmerge(left, [2, 3, 6, 7], 0, 0, 3, 4)
    assert k == 6 and left[:6] == [1, 2, 3, 5, 6, 7] and left[6] == 0
    # a value repeated inside one list stays repeated; one shared copy merges
    left = [4, 4, 9, 0, 0]
    k = cp.lmerge(left, [4, 5], 0, 0, 3, 2)
    assert k == 4 and left[:4] == [4, 4, 5, 9]
    # offsets address segments of both lists; the rest of left is untouched
    left = np.array([100, 2, 4, 8, 0, 0, 0, -1], dtype=np.int64)
    right = np.array([7, 7, 1, 4, 9], dtype=np.int64)
    k = cp.lmerge(left, right, 1, 2, 3, 3)
    assert k == 5
    assert left.tolist() == [100, 1, 2, 4, 8, 9, 0, -1]
    # empty segments
    left = [3, 5]
    assert cp.lmerge(left, [], 0, 0, 2, 0) == 2 and left == [3, 5]
    left = [0, 0]
    assert cp.lmerge(left, [6, 7], 0, 0, 0, 2) == 2 and left == [6, 7]
    # default lengths: the whole lists
    left = np.array([1, 5, 0, 0], dtype=np.int64)
    with pytest.raises(ValueError):
        cp.lmerge(left, np.array([2], dtype=np.int64))  # 0 and 0 would merge to 5 entries
    # left too short for the merged result
    with pytest.raises(ValueError):
        cp.lmerge([1, 2], [3, 4], 0, 0, 2, 2)
    with pytest.raises(ValueError):
        cp.lmerge([1, 2], [3], 0, 0, 5, 1)


def test_lmerge_random_against_multiset_union():
    rng = np.random.default_rng(11)
    for _ in range(300):
        a = sorted(rng.integers(0, 12, rng.integers(0, 9)).tolist())
        b = sorted(rng.integers(0, 12, rng.integers(0, 9)).tolist())
        pad = rng.integers(0, 3)
        left = [-5] * pad + a + [-7] * (len(b) + 2)
        right = [99] + b
        k = cp.lmerge(left, right, pad, 1, len(a), len(b))
        want = _multiset_max_union(a, b)
        assert k == len(want)
        assert left[pad:pad + k] == want
        assert left[:pad] == [-5] * pad
        assert left[pad + k:] == [-7] * (len(a) + len(b) + 2 - k)


def test_lmerge_sets_match_embed_union():
    # on duplicate-free lists (the symbolic layer's use) it is the sorted union
    rng = np.random.default_rng(3)
    for _ in range(50):
        a = np.unique(rng.integers(0, 40, 15))
        b = np.unique(rng.integers(0, 40, 15))
        left = np.zeros(a.size + b.size, dtype=np.int64)
        left[:a.size] = a
        k = cp.lmerge(left, b, 0, 0, a.size, b.size)
        assert np.array_equal(left[:k], np.union1d(a, b))


# ---------------------------------------------- cvxopt-style arguments
class _Buf(object):
    """A 1-D integer or float buffer like a cvxopt matrix column: supports
    the buffer protocol through array.array."""

    def __init__(self, vals, code):
        self._a = array.array(code, vals)

    def __buffer__(self, flags):  # Python >= 3.12
        return memoryview(self._a)

    def __array__(self, dtype=None, copy=None):
        a = np.frombuffer(self._a, dtype=np.float64 if self._a.typecode == "d" else np.int64)
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return len(self._a)


class CCSMatrix(object):
    """Minimal cvxopt-style spmatrix exposing ``size`` and ``CCS``."""

    def __init__(self, A):
        A = sp.csc_matrix(A)
        A.sort_indices()
        self.size = A.shape
        self.CCS = (_Buf(A.indptr.tolist(), "q"), _Buf(A.indices.tolist(), "q"),
                    _Buf(A.data.tolist(), "d"))


class IJVMatrix(object):
    """Minimal cvxopt-style spmatrix exposing ``size`` and ``I``, ``J``, ``V``."""

    def __init__(self, A):
        A = sp.coo_matrix(A)
        self.size = A.shape
        self.I = _Buf(A.row.astype(np.int64).tolist(), "q")
        self.J = _Buf(A.col.astype(np.int64).tolist(), "q")
        self.V = _Buf(A.data.tolist(), "d")


SYMB_ATTRS = ["n", "Nsn", "snpost", "snptr", "snode", "snpar", "relptr", "relidx", "chptr",
              "chidx", "blkptr", "p", "ip", "sncolptr", "snrowidx", "clique_number"]


@pytest.mark.parametrize("wrap", [CCSMatrix, IJVMatrix])
@pytest.mark.parametrize("make", [A17, A23, lambda: problems.grid2d(12, seed=0)])
def test_symbolic_and_assembly_from_cvxopt_style_sparse(wrap, make):
    A = make()
    ref = cp.symbolic(A, p=cp.amd.order)
    W = wrap(A)
    assert misc.is_sparse(W)
    s = cp.symbolic(W, p=cp.amd.order)
    for name in SYMB_ATTRS:
        a, b = getattr(ref, name), getattr(s, name)
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            assert np.array_equal(np.asarray(a), np.asarray(b)), name
        else:
            assert a == b, name
    assert ref.memory == s.memory
    X = cp.cspmatrix(ref, device="cpu") + A
    Y = cp.cspmatrix(ref, device="cpu") + wrap(A)
    assert torch.equal(X.blkval, Y.blkval)
    Y -= wrap(A)
    assert float(Y.blkval.abs().max()) == 0.0
    # tril / symmetrize / perm / peo / maxcardsearch accept it too
    assert (misc.tril(W) != misc.tril(A)).nnz == 0
    assert (cp.symmetrize(W) != cp.symmetrize(A)).nnz == 0
    assert cp.peo(W, np.arange(A.shape[0])) == cp.peo(A, np.arange(A.shape[0]))
    assert np.array_equal(cp.maxcardsearch(W), cp.maxcardsearch(A))


def test_as_sparse_rejects_dense_and_complex():
    with pytest.raises(TypeError):
        misc.as_sparse(np.eye(3))

    class Z(object):
        size = (2, 2)
        I, J, V = [0, 1], [0, 1], np.array([1j, 2.0])

    with pytest.raises(TypeError):
        misc.as_sparse(Z())


def test_host_dense_views_share_memory():
    a = np.asfortranarray(np.arange(12.0).reshape(3, 4))
    m = memoryview(a)                       # the buffer protocol, as a cvxopt matrix
    h = numeric.host_dense(m)
    assert h.shape == (3, 4) and np.shares_memory(h, a)
    h[1, 2] = -1.0
    assert a[1, 2] == -1.0
    # Fortran-order flattening is the column-major storage sequence
    assert np.array_equal(h.reshape(-1, order="F"), a.ravel(order="F"))
    with pytest.raises(TypeError):
        numeric.host_dense(np.zeros((2, 2), dtype=np.int64))
    ro = np.zeros(3)
    ro.flags.writeable = False
    with pytest.raises(ValueError):
        numeric.host_dense(ro)


# ------------------------------------------- strided GPU-side operands
def _check_colmajor(T):
    work, ld, avail, copied = numeric.colmajor_work(T)
    rows = T.shape[0]
    cols = 1 if T.dim() == 1 else T.shape[1]
    assert ld >= max(rows, 1)
    # element (i, j) of T sits at offset i + j*ld from work's first element
    flat = torch.as_strided(work, (work.untyped_storage().nbytes() // 8 - work.storage_offset(),),
                            (1,), work.storage_offset())
    assert (cols - 1) * ld + rows <= avail <= flat.numel()
    for j in range(cols):
        col = T if T.dim() == 1 else T[:, j]
        assert torch.equal(flat[j * ld:j * ld + rows], col)
    return copied


def test_colmajor_work_views():
    M = torch.arange(60, dtype=torch.float64).reshape(10, 6)        # row-major
    assert _check_colmajor(M) is True
    assert _check_colmajor(M.t().contiguous().t()) is False         # column-major
    assert _check_colmajor(M[:, 0]) is True                         # strided column
    assert _check_colmajor(M[:, 2:3]) is True                       # (n, 1), row stride 6
    F = M.t().contiguous().t()
    assert _check_colmajor(F[:, 1:4]) is False                      # column-major slice
    assert _check_colmajor(F[2:7, ::2]) is False                    # ld = 10 >= 5 rows
    assert _check_colmajor(torch.arange(5.0, dtype=torch.float64)) is False
    e = torch.ones(4, 1, dtype=torch.float64).expand(4, 3)          # stride (1, 0)
    assert _check_colmajor(e) is True
    s = torch.zeros(8, dtype=torch.float64)[::2]                    # 1-D stride 2
    assert _check_colmajor(s) is True
