"""Plain-Python restatement of CHOMPACK's chordal orderings.

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/; the product
(chompack_amd) never imports it.  maxcardsearch follows mcs.py:4-57 and
maxchord follows maxchord.py:5-83 line by line, with the CCS arrays of
symmetrize(A) (rows ascending per column, as cvxopt's CCS) in place of
cvxopt matrices.  Pinned in tests/test_orderings.py by the reference
documentation's 17 x 17 example (a perfect elimination ordering without
fill), by chordality and brute-force maximality of maxchord's subgraph, and
by maxchord(A) = A for chordal A.
"""
import numpy as np
import scipy.sparse as sp

from chompack_amd.misc import symmetrize

__all__ = ["maxcardsearch", "maxchord"]


def _ccs(A):
    As = sp.csc_matrix(symmetrize(A))
    As.sort_indices()
    return As.indptr, As.indices, As.data


def maxcardsearch(A, ve=None):
    n = A.shape[0]
    if ve is None:
        ve = n - 1
    cp, ri, _ = _ccs(A)
    p = np.zeros(n, dtype=np.int64)
    w = [0] * n
    max_w = 0
    S = [list(range(ve)) + list(range(ve + 1, n)) + [ve]] + [[] for _ in range(n - 1)]
    for i in range(n - 1, -1, -1):
        while True:
            if len(S[max_w]) > 0:
                v = S[max_w].pop()
                if w[v] >= 0:
                    break
            else:
                max_w -= 1
        p[i] = v
        w[v] = -1
        for r in ri[cp[v]:cp[v + 1]]:
            if w[r] >= 0:
                w[r] += 1
                S[w[r]].append(r)
                max_w = max(max_w, w[r])
    return p


def maxchord(A, ve=None):
    n = A.shape[0]
    if ve is None:
        ve = n - 1
    cp, ri, val = _ccs(A)
    p = np.zeros(n, dtype=np.int64)
    w = [0] * n
    max_w = 0
    S = [list(range(ve)) + list(range(ve + 1, n)) + [ve]] + [[] for _ in range(n - 1)]
    C = [set() for _ in range(n)]
    E = [[] for _ in range(n)]
    V = [[] for _ in range(n)]
    for i in range(n - 1, -1, -1):
        while True:
            if len(S[max_w]) > 0:
                v = S[max_w].pop()
                if w[v] >= 0:
                    break
            else:
                max_w -= 1
        p[i] = v
        w[v] = -1
        for ii in range(cp[v], cp[v + 1]):
            u = ri[ii]
            d = val[ii]
            if w[u] >= 0:
                if C[u].issubset(C[v]):
                    C[u].update([v])
                    w[u] += 1
                    S[w[u]].append(u)
                    max_w = max(max_w, w[u])
                    E[min(u, v)].append(max(u, v))
                    V[min(u, v)].append(d)
            elif u == v:
                E[u].append(u)
                V[u].append(d)
    rows = [r for Ej in E for r in Ej]
    cols = [j for j, Ej in enumerate(E) for _ in Ej]
    vals = [d for Vj in V for d in Vj]
    Am = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    return Am, p
