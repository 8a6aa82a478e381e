"""Symbolic factorization (CPU, integer, bit-exact) and the chordal matrix type.

Drop-in mirror of chompack's ``symbolic`` / ``cspmatrix`` classes
(src/python/symbolic.py:499-1210).  The integer algorithms (elimination tree,
postorder, column counts, Pothen-Sun supernodes, embedding, relative indices)
run in the host part of libchompack_amd.so (csrc/symbolic.cpp); amalgamation,
renumbering and the memory plan are vectorised NumPy here.  Every attribute
the reference's C wrappers read (cbase.c:194-246, 972-1041) is exposed with
the same meaning.

``cspmatrix.blkval`` is a 1-D float64 ``torch`` tensor, normally resident in
HBM (device ``cuda``); the block layout is the reference's: supernode k owns a
column-major ``nj x nn`` block at ``blkptr[k]`` with leading dimension ``nj``,
rows ``snrowidx[sncolptr[k]:sncolptr[k+1]]`` (only the lower trapezoid is
meaningful).
"""
import numpy as np
import scipy.sparse as sp

from . import _native
from .misc import coo_triplets, symmetrize, perm, tril, as_sparse, is_sparse

__all__ = ["symbolic", "cspmatrix", "merge_size_fill", "peo", "post_order",
           "etree", "counts"]


def _csc_pattern(A):
    """(colptr, rowind) int64 arrays of a sparse matrix, rows sorted."""
    A = sp.csc_matrix(as_sparse(A))
    A.sort_indices()
    return _native.as_i64(A.indptr), _native.as_i64(A.indices)


def etree(A):
    """Elimination tree of the upper triangle of A (symbolic.py:49-72)."""
    n = A.shape[0]
    cp, ri = _csc_pattern(A)
    par = np.empty(n, dtype=np.int64)
    _native.lib().chp_etree(n, _native.ptr(cp), _native.ptr(ri), _native.ptr(par))
    return par


def post_order(parent):
    """Postorder of a forest, roots have parent[j] == j (symbolic.py:26-47)."""
    parent = _native.as_i64(parent)
    n = len(parent)
    post = np.empty(n, dtype=np.int64)
    _native.lib().chp_post_order(n, _native.ptr(parent), _native.ptr(post))
    return post


def counts(A, parent, post):
    """Column counts of the Cholesky factor (symbolic.py:95-131)."""
    n = A.shape[0]
    cp, ri = _csc_pattern(A)
    parent = _native.as_i64(parent)
    post = _native.as_i64(post)
    cc = np.empty(n, dtype=np.int64)
    _native.lib().chp_counts(n, _native.ptr(cp), _native.ptr(ri), _native.ptr(parent),
                             _native.ptr(post), _native.ptr(cc))
    return cc


def _supernodes(par, post, colcount):
    n = len(par)
    snode = np.empty(n, dtype=np.int64)
    snptr = np.empty(n + 1, dtype=np.int64)
    snpar = np.empty(n, dtype=np.int64)
    nsn = _native.lib().chp_supernodes(n, _native.ptr(par), _native.ptr(post),
                                       _native.ptr(colcount), _native.ptr(snode),
                                       _native.ptr(snptr), _native.ptr(snpar))
    return snode, snptr[:nsn + 1].copy(), snpar[:nsn].copy()


def _amalgamate(colcount, snode, snptr, snpar, snpost, merge_function):
    """Greedy supernodal amalgamation in supernodal postorder
    (symbolic.py:246-326)."""
    N = len(snpost)
    ch = {}
    for j in snpost:
        ch.setdefault(int(snpar[j]), []).append(int(j))
    snlist = [list(snode[snptr[k]:snptr[k + 1]]) for k in range(N)]
    snpar_ = snpar.copy()
    colcount_ = colcount.copy()
    for k in snpost:
        k = int(k)
        if snpar_[k] != k:
            p = int(snpar_[k])
            colk = int(colcount_[snlist[k][0]])
            colp = int(colcount_[snlist[p][0]])
            nk = len(snlist[k])
            np_ = len(snlist[p])
            if merge_function and merge_function(colp, colk, np_, nk):
                snlist[p] = sorted(snlist[k] + snlist[p])
                snlist[k] = None
                colcount_[snlist[p][0]] = colp + nk
                if k in ch:
                    for c in ch[k]:
                        snpar_[c] = p
                    ch.setdefault(p, [])
                    ch[p] += ch[k]
                snpar_[k] = k
    L = [i for i, s in enumerate(snlist) if s is not None]
    snptr_ = np.zeros(len(L) + 1, dtype=np.int64)
    snode_ = snode.copy()
    for i, l in enumerate(L):
        snptr_[i + 1] = snptr_[i] + len(snlist[l])
        snode_[snptr_[i]:snptr_[i + 1]] = snlist[l]
    pos = {l: i for i, l in enumerate(L)}
    snpar_ = np.array([pos[int(snpar_[l])] for l in L], dtype=np.int64)
    snpost_ = post_order(snpar_)
    return colcount_, snode_, snptr_, snpar_, snpost_


def merge_size_fill(tsize=8, tfill=8):
    """Merge heuristic (symbolic.py:451-497): merge supernode k into its
    parent if both have at most ``tsize`` nodes or the merge adds at most
    ``tfill`` lower-triangular nonzeros."""
    assert tsize >= 0, "tsize must be nonnegative"
    assert tfill >= 0, "tfill must be nonnegative"

    def fmerge(ccp, cck, np_, nk):
        fill = (ccp - (cck - nk)) * nk
        return bool(fill <= tfill or (nk <= tsize and np_ <= tsize))
    fmerge.tsize = tsize
    fmerge.tfill = tfill
    return fmerge


def peo(A, p):
    """True if p is a perfect elimination ordering of the lower triangle of A
    (symbolic.py:413-449)."""
    assert is_sparse(A), "A must be a sparse matrix"
    A = as_sparse(A)
    n = A.shape[0]
    assert A.shape[1] == n, "A must be a square matrix"
    p = np.asarray(p, dtype=np.int64).ravel()
    assert len(p) == n, "length of p must be equal to the order of A"
    As = sp.csc_matrix(symmetrize(A))
    As.sort_indices()
    cp, ri = As.indptr, As.indices
    ip = np.empty(n, dtype=np.int64)
    ip[p] = np.arange(n)
    for k in range(n):
        v = p[k]
        r = set(int(x) for x in ri[cp[v]:cp[v + 1]] if ip[x] > k)
        for rj in r:
            if not r.issubset(set(int(x) for x in ri[cp[rj]:cp[rj + 1]])):
                return False
    return True


class symbolic(object):
    """Symbolic factorization object (symbolic.py:499-890).

    ``symbolic(A, p=None, merge_function=None, supernodal=True)`` where A is a
    square sparse matrix (``scipy.sparse``, or a cvxopt-style ``spmatrix``
    exposing ``size`` and ``CCS`` or ``I``/``J``/``V``; only its lower
    triangle is read), ``p`` a
    permutation vector or an ordering routine (e.g. ``chompack_amd.amd.order``),
    and ``merge_function`` an amalgamation heuristic such as
    ``merge_size_fill(tsize, tfill)``.
    """

    def __init__(self, A, p=None, merge_function=None, **kwargs):
        assert is_sparse(A), "A must be a sparse matrix"
        A = as_sparse(A)
        assert A.shape[0] == A.shape[1], "A must be a square matrix"
        supernodal = kwargs.get("supernodal", True)
        n = A.shape[0]

        Ap = sp.csc_matrix(symmetrize(A))
        Ap.sort_indices()
        nnz_Ap = (Ap.nnz + n) // 2
        dI, dJ, _ = coo_triplets(Ap)
        assert int(np.count_nonzero(dI == dJ)) == n, \
            "the sparsity pattern of A must include diagonal elements"

        if p is not None:
            if callable(p):
                p = p(Ap)
            p = np.asarray(p, dtype=np.int64).ravel()
            assert len(p) == n, "length of permutation vector must be equal to the order of A"
            Ap = sp.csc_matrix(perm(Ap, p))
            Ap.sort_indices()

        cp, ri = _native.as_i64(Ap.indptr), _native.as_i64(Ap.indices)
        L = _native.lib()
        par = np.empty(n, dtype=np.int64)
        L.chp_etree(n, _native.ptr(cp), _native.ptr(ri), _native.ptr(par))
        post = post_order(par)
        colcount = np.empty(n, dtype=np.int64)
        L.chp_counts(n, _native.ptr(cp), _native.ptr(ri), _native.ptr(par),
                     _native.ptr(post), _native.ptr(colcount))
        nnz_Ae = int(colcount.sum())
        if supernodal:
            snode, snptr, snpar = _supernodes(par, post, colcount)
            snpost = post_order(snpar)
        else:
            snpar, snpost = par, post
            snode = np.arange(n, dtype=np.int64)
            snptr = np.arange(n + 1, dtype=np.int64)
        if merge_function:
            colcount, snode, snptr, snpar, snpost = _amalgamate(
                colcount, snode, snptr, snpar, snpost, merge_function)

        # renumber: supernodes contiguous and postordered (symbolic.py:564-584)
        Nsn = len(snpar)
        sizes = np.diff(snptr)
        pp = np.concatenate([snode[snptr[k]:snptr[k + 1]] for k in snpost]) \
            if Nsn else np.zeros(0, dtype=np.int64)
        snptr2 = np.zeros(Nsn + 1, dtype=np.int64)
        np.cumsum(sizes[snpost], out=snptr2[1:])
        colcount = colcount[pp]
        snposti = np.empty(Nsn, dtype=np.int64)
        snposti[snpost] = np.arange(Nsn, dtype=np.int64)
        snpar = snposti[snpar[snpost]]
        snode = np.arange(n, dtype=np.int64)
        snpost = np.arange(Nsn, dtype=np.int64)
        snptr = snptr2

        Ap = sp.csc_matrix(perm(Ap, pp))
        Ap.sort_indices()
        self._p = pp if p is None else p[pp]
        self._ip = np.empty(n, dtype=np.int64)
        self._ip[self._p] = np.arange(n, dtype=np.int64)

        # embedding and relative indices (symbolic.py:328-411)
        Alo = sp.csc_matrix(tril(Ap))
        Alo.sort_indices()
        cpl, ril = _native.as_i64(Alo.indptr), _native.as_i64(Alo.indices)
        sncolptr = np.zeros(Nsn + 1, dtype=np.int64)
        np.cumsum(colcount[snptr[:-1]], out=sncolptr[1:])
        snrowidx = np.full(int(sncolptr[-1]), -1, dtype=np.int64)
        rc = L.chp_embed(Nsn, _native.ptr(cpl), _native.ptr(ril), _native.ptr(snptr),
                         _native.ptr(snpar), _native.ptr(sncolptr), _native.ptr(snrowidx))
        if rc != 0:
            raise ArithmeticError("symbolic embedding failed at supernode %d" % (-1 - rc))
        nnk = np.diff(snptr)
        nak = np.diff(sncolptr) - nnk
        relptr = np.zeros(Nsn + 1, dtype=np.int64)
        np.cumsum(np.where(snpar != np.arange(Nsn), nak, 0), out=relptr[1:])
        relidx = np.full(int(relptr[-1]), -1, dtype=np.int64)
        rc = L.chp_relative_idx(Nsn, _native.ptr(snptr), _native.ptr(snpar),
                                _native.ptr(sncolptr), _native.ptr(snrowidx),
                                _native.ptr(relptr), _native.ptr(relidx))
        if rc != 0:
            raise ArithmeticError("relative indices failed at supernode %d" % (-1 - rc))

        # children lists in postorder (symbolic.py:590-604)
        nonroot = snpar != np.arange(Nsn)
        chptr = np.zeros(Nsn + 1, dtype=np.int64)
        np.cumsum(np.bincount(snpar[nonroot], minlength=Nsn), out=chptr[1:])
        kids = np.nonzero(nonroot)[0]
        chidx = kids[np.argsort(snpar[kids], kind="stable")].astype(np.int64)

        blkptr = np.zeros(Nsn + 1, dtype=np.int64)
        np.cumsum(nnk * (nnk + nak), out=blkptr[1:])

        # memory plan of the reference's sequential sweep (symbolic.py:606-680):
        # its LIFO stack holds supernode j's update matrix from step j (after
        # its push) until the step of its parent (which pops it before its
        # own push), so the stack after step k holds exactly the j <= k whose
        # parent is > k.  A difference array over those intervals gives the
        # stack's size, depth and element count after every step at once.
        na_a = (relptr[1:] - relptr[:-1]).astype(np.int64)
        nj_a = nnk.astype(np.int64) + na_a
        live = np.nonzero(na_a > 0)[0]           # the pushed (non-root) supernodes
        stats = {}
        for key, wgt in (("stack_mem", na_a * na_a), ("stack_depth", np.ones(Nsn, np.int64)),
                         ("stack_solve", na_a)):
            d = np.zeros(Nsn + 1, dtype=np.int64)
            np.add.at(d, live, wgt[live])
            np.add.at(d, snpar[live], -wgt[live])
            stats[key] = int(np.cumsum(d[:Nsn]).max()) if Nsn else 0
        cln = int(nj_a.max()) if Nsn else 0
        stack_depth, stack_mem, stack_solve = (stats["stack_depth"], stats["stack_mem"],
                                               stats["stack_solve"])
        self.frontal_len = cln * cln
        self.stack_len = stack_mem
        self.stack_size = stack_depth

        self._clique_number = int(cln)
        self._n = int(n)
        self._Nsn = int(Nsn)
        self._snode = snode
        self._snptr = snptr
        self._chptr = chptr
        self._chidx = chidx
        self._snpar = snpar
        self._snpost = snpost
        self._relptr = relptr
        self._relidx = relidx
        self._sncolptr = sncolptr
        self._snrowidx = snrowidx
        self._blkptr = blkptr
        self._fill = (nnz_Ae - nnz_Ap, self.nnz - nnz_Ae)
        self._memory = {"stack_depth": int(stack_depth), "stack_mem": int(stack_mem),
                        "frontal_mem": int(cln) ** 2, "stack_solve": int(stack_solve)}
        self._plans = {}
        self._export_cache = None

    # ---------------------------------------------------------------- views
    def __repr__(self):
        return "<symbolic factorization, n=%i, nnz=%i, nsn=%i>" % (self.n, self.nnz, self.Nsn)

    def __str__(self):
        S = self.sparsity_pattern(reordered=True, symmetric=True).toarray()
        rows = []
        for i in range(self.n):
            rows.append("[" + " ".join("X" if S[i, j] != 0 else " " for j in range(self.n)) + "]")
        return "\n".join(rows) + "\n"

    def sparsity_pattern(self, reordered=True, symmetric=True):
        """Filled pattern as a sparse matrix of ones (symbolic.py:692-703)."""
        X = cspmatrix(self, 1.0, device="cpu")
        return X.spmatrix(reordered=reordered, symmetric=symmetric)

    n = property(lambda self: self._n, doc="Number of nodes (matrix order)")
    Nsn = property(lambda self: self._Nsn, doc="Number of supernodes")
    snode = property(lambda self: self._snode)
    snptr = property(lambda self: self._snptr)
    chptr = property(lambda self: self._chptr)
    chidx = property(lambda self: self._chidx)
    snpar = property(lambda self: self._snpar)
    snpost = property(lambda self: self._snpost)
    relptr = property(lambda self: self._relptr)
    relidx = property(lambda self: self._relidx)
    sncolptr = property(lambda self: self._sncolptr)
    snrowidx = property(lambda self: self._snrowidx)
    blkptr = property(lambda self: self._blkptr)
    fill = property(lambda self: self._fill)
    p = property(lambda self: self._p)
    ip = property(lambda self: self._ip)
    memory = property(lambda self: self._memory)
    clique_number = property(lambda self: self._clique_number)

    @property
    def nnz(self):
        """Number of lower-triangular nonzeros (symbolic.py:818-828)."""
        nn = np.diff(self._snptr)
        na = np.diff(self._relptr)
        return int(np.sum(nn * (nn + 1) // 2 + nn * na))

    def cliques(self, reordered=True):
        out = [list(self._snrowidx[self._sncolptr[k]:self._sncolptr[k + 1]]) for k in range(self.Nsn)]
        if reordered:
            return [[int(v) for v in c] for c in out]
        return [[int(self._p[v]) for v in c] for c in out]

    def separators(self, reordered=True):
        out = []
        for k in range(self.Nsn):
            nn = self._snptr[k + 1] - self._snptr[k]
            s = self._snrowidx[self._sncolptr[k] + nn:self._sncolptr[k + 1]]
            out.append([int(v) if reordered else int(self._p[v]) for v in s])
        return out

    def supernodes(self, reordered=True):
        out = []
        for k in range(self.Nsn):
            s = self._snode[self._snptr[k]:self._snptr[k + 1]]
            out.append([int(v) if reordered else int(self._p[v]) for v in s])
        return out

    def parent(self):
        """Supernodal parent list; a root k has parent k (symbolic.py:884-890)."""
        return [int(v) for v in self._snpar]

    # ------------------------------------------------- cached export indices
    def _lower_map(self):
        """For every stored lower-trapezoid entry: (blkval offset, row, col) in
        the reordered numbering.  Cached; used by assembly and export."""
        if self._export_cache is None:
            nn = np.diff(self._snptr)
            nj = np.diff(self._sncolptr)
            offs, rows, cols = [], [], []
            # group supernodes by (nn, nj) shape is overkill; per supernode numpy
            for k in range(self.Nsn):
                a, b = int(nn[k]), int(nj[k])
                c0 = int(self._snptr[k])
                rr = self._snrowidx[self._sncolptr[k]:self._sncolptr[k + 1]]
                ii, jj = np.tril_indices(b, 0, a)  # rows ii >= jj, jj < a
                offs.append(self._blkptr[k] + jj * b + ii)
                rows.append(rr[ii])
                cols.append(c0 + jj)
            if offs:
                self._export_cache = (np.concatenate(offs), np.concatenate(rows),
                                      np.concatenate(cols))
            else:
                z = np.zeros(0, dtype=np.int64)
                self._export_cache = (z, z, z)
        return self._export_cache

    def _plan(self, device):
        """Device plan (level schedule + index arrays in HBM), built once per
        device and cached on this object."""
        from . import numeric
        key = str(device)
        if key not in self._plans:
            self._plans[key] = numeric.Plan(self, device)
        return self._plans[key]


class cspmatrix(object):
    """Chordal sparse matrix (symbolic.py:892-1210).

    ``cspmatrix(symb, blkval=None, factor=False, device=None)``: ``blkval`` is
    used as is (not copied) when it is a float64 torch tensor of length
    ``symb.blkptr[-1]``; a float fills the array with that value; None fills
    it with zeros.  ``device`` defaults to ``cuda`` when a GPU is present.
    """

    def __init__(self, symb, blkval=None, factor=False, device=None):
        import torch
        assert isinstance(symb, symbolic), "symb must be an instance of symbolic"
        self._is_factor = factor
        self.symb = symb
        N = int(symb.blkptr[-1])
        if device is None:
            device = blkval.device if isinstance(blkval, torch.Tensor) else \
                ("cuda" if torch.cuda.is_available() else "cpu")
        if blkval is None:
            self.blkval = torch.zeros(N, dtype=torch.float64, device=device)
        elif isinstance(blkval, float):
            self.blkval = torch.full((N,), blkval, dtype=torch.float64, device=device)
        else:
            if not isinstance(blkval, torch.Tensor):
                blkval = torch.as_tensor(np.asarray(blkval, dtype=np.float64), device=device)
            assert blkval.numel() == N, "dimension mismatch: length of blkval is incorrect"
            assert blkval.dtype == torch.float64
            self.blkval = blkval

    def __repr__(self):
        kind = "chordal sparse matrix (factor)" if self.is_factor else "chordal sparse matrix"
        return "<%ix%i %s, tc='d', nnz=%i, nsn=%i>" % (self.symb.n, self.symb.n, kind,
                                                       self.symb.nnz, self.symb.Nsn)

    def __str__(self):
        return str(self.spmatrix(reordered=True, symmetric=False).toarray())

    @property
    def is_factor(self):
        return self._is_factor

    @is_factor.setter
    def is_factor(self, value):
        self._is_factor = value

    @property
    def size(self):
        return (self.symb.n, self.symb.n)

    # ------------------------------------------------------------ arithmetic
    def __iadd__(self, B):
        assert self.is_factor is False, "Addition of cspmatrix factors not supported"
        if isinstance(B, cspmatrix):
            assert self.symb is B.symb, "Symbolic factorization mismatch"
            assert B.is_factor is False, "Addition of cspmatrix factors not supported"
            self.blkval.add_(B.blkval)
        elif is_sparse(B):
            self._iadd_spmatrix(as_sparse(B))
        else:
            raise TypeError
        return self

    def __isub__(self, B):
        assert self.is_factor is False, "Addition of cspmatrix factors not supported"
        if isinstance(B, cspmatrix):
            assert self.symb is B.symb, "Symbolic factorization mismatch"
            assert B.is_factor is False, "Addition of cspmatrix factors not supported"
            self.blkval.sub_(B.blkval)
        elif is_sparse(B):
            self._iadd_spmatrix(as_sparse(B), alpha=-1.0)
        else:
            raise TypeError
        return self

    def __add__(self, B):
        C = self.copy()
        C.__iadd__(B)
        return C

    def __sub__(self, B):
        C = self.copy()
        C.__isub__(B)
        return C

    def __imul__(self, a):
        if isinstance(a, (float, int)):
            self.blkval.mul_(float(a))
        else:
            raise NotImplementedError("only scalar multiplication has been implemented")
        return self

    def __mul__(self, a):
        ret = self.copy()
        ret.__imul__(a)
        return ret

    def __rmul__(self, a):
        return self.__mul__(a)

    def copy(self):
        """New cspmatrix sharing ``symb`` with a copy of ``blkval``."""
        b = self.blkval
        if getattr(b, "is_cuda", False) and str(b.dtype) == "torch.float64" \
                and b.is_contiguous() and b.data_ptr() % 16 == 0:
            from . import numeric
            c = b.new_empty(b.shape)
            numeric.copy_blkval(c, b)
            return cspmatrix(self.symb, blkval=c, factor=self.is_factor)
        return cspmatrix(self.symb, blkval=b.clone(), factor=self.is_factor)

    # ------------------------------------------------------ assembly/export
    def _entry_offsets(self, I, J):
        """blkval offsets of reordered lower-triangular entries (I >= J)."""
        symb = self.symb
        n = symb.n
        k = np.searchsorted(symb.snptr, J, side="right") - 1
        keys_all = np.repeat(np.arange(symb.Nsn, dtype=np.int64), np.diff(symb.sncolptr)) * n \
            + symb.snrowidx
        g = np.searchsorted(keys_all, k * n + I)
        ok = (g < len(keys_all))
        ok[ok] = keys_all[g[ok]] == (k * n + I)[ok]
        if not np.all(ok):
            raise ValueError("entry outside the chordal sparsity pattern")
        pos = g - symb.sncolptr[k]
        nj = symb.sncolptr[k + 1] - symb.sncolptr[k]
        return symb.blkptr[k] + nj * (J - symb.snptr[k]) + pos

    def _iadd_spmatrix(self, X, alpha=1.0):
        """Add tril(P sym(X) P^T) to the blocks (symbolic.py:1138-1174)."""
        import torch
        assert self.is_factor is False, "cannot add spmatrix to a cspmatrix factor"
        symb = self.symb
        if symb.p is not None:
            Xp = tril(perm(symmetrize(X), symb.p))
        else:
            Xp = tril(X)
        I, J, V = coo_triplets(Xp)
        off = self._entry_offsets(I, J)
        vals = torch.as_tensor(alpha * V, dtype=torch.float64, device=self.blkval.device)
        idx = torch.as_tensor(off, device=self.blkval.device)
        self.blkval.index_add_(0, idx, vals)

    def add_projection(self, A, alpha=1.0, beta=1.0, reordered=False):
        """X := alpha*proj(A) + beta*X for a dense symmetric A
        (symbolic.py:1176-1210)."""
        import torch
        assert self.is_factor is False, "cannot project matrix onto a cspmatrix factor"
        A = np.asarray(A, dtype=np.float64)
        symb = self.symb
        if symb.p is not None and reordered is False:
            A = np.tril(A)
            A = A + A.T
            A[np.diag_indices_from(A)] *= 0.5
            A = A[np.ix_(symb.p, symb.p)]
        out = np.empty(int(symb.blkptr[-1]))
        nn = np.diff(symb.snptr)
        for k in range(symb.Nsn):
            rr = symb.snrowidx[symb.sncolptr[k]:symb.sncolptr[k + 1]]
            cc = symb.snode[symb.snptr[k]:symb.snptr[k + 1]]
            out[symb.blkptr[k]:symb.blkptr[k + 1]] = A[np.ix_(rr, cc)].ravel(order="F")
        new = torch.as_tensor(out, device=self.blkval.device)
        self.blkval.mul_(beta).add_(alpha * new)
        del nn

    def spmatrix(self, reordered=True, symmetric=False):
        """Export to a scipy.sparse matrix (symbolic.py:1032-1111).  For a
        factor, the (2,1) blocks are multiplied back by L11 so the result is
        the lower-triangular factor L of P A P^T."""
        symb = self.symb
        b = self.blkval
        if self.is_factor:
            if symmetric:
                raise ValueError("'symmetric = True' not implemented for Cholesky factors")
            if not reordered:
                raise ValueError("'reordered = False' not implemented for Cholesky factors")
        if self.is_factor and getattr(b, "is_cuda", False):
            # L21 = B21 L11 for every supernode on the GPU (k_b21_l11)
            from . import numeric, _native
            blk = numeric._check_blkval(self)
            out = blk.new_empty(blk.shape)
            plan = symb._plan(blk.device)
            rc = _native.lib().chp_factor_export(plan.handle, blk.data_ptr(), out.data_ptr(),
                                                 numeric._stream(blk))
            numeric._raise_rc(rc, "spmatrix")
            vals = out.cpu().numpy()
        else:
            vals = b.detach().cpu().numpy().copy() if hasattr(b, "detach") else np.array(b, dtype=float)
        if self.is_factor and not getattr(b, "is_cuda", False):
            # a host cspmatrix: the same products with numpy, per supernode
            nnk = np.diff(symb.snptr)
            njk = np.diff(symb.sncolptr)
            for k in range(symb.Nsn):
                nn, nj = int(nnk[k]), int(njk[k])
                na = nj - nn
                if na == 0:
                    continue
                blk = vals[symb.blkptr[k]:symb.blkptr[k + 1]].reshape((nn, nj)).T
                L11 = np.tril(blk[:nn, :])
                blk[nn:, :] = blk[nn:, :] @ L11
        off, rows, cols = symb._lower_map()
        v = vals[off]
        n = symb.n
        T = sp.csc_matrix((v, (rows, cols)), shape=(n, n))
        if reordered:
            if not symmetric:
                return T
            return symmetrize(T)
        S = perm(symmetrize(T), symb.ip)
        return S if symmetric else tril(S)

    def diag(self, reordered=True):
        """Diagonal entries as a float64 numpy vector (symbolic.py:1113-1128)."""
        symb = self.symb
        idx = self._diag_offsets()
        D = self.blkval.detach()[idx].cpu().numpy() if hasattr(self.blkval, "detach") \
            else np.asarray(self.blkval)[idx]
        if reordered:
            return D
        return D[symb.ip]

    def _diag_offsets(self):
        symb = self.symb
        if not hasattr(symb, "_diag_cache"):
            nn = np.diff(symb.snptr)
            nj = np.diff(symb.sncolptr)
            k = np.repeat(np.arange(symb.Nsn), nn)
            i = np.arange(symb.n) - symb.snptr[k]
            symb._diag_cache = symb.blkptr[k] + i * nj[k] + i
        import torch
        return torch.as_tensor(symb._diag_cache, device=self.blkval.device)
