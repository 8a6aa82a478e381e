"""Synthetic SPD test problems (grids, batches of grids, random geometric
graphs) with the value rule used throughout the benchmarks and tests.

Value rule: for every undirected edge (i, j), i > j, a_ij = -U(0,1); each
diagonal entry is the row's sum of |off-diagonals| plus 1.  The matrix is SPD
with Gershgorin bounds [1, 2*maxdeg + 1].  Random numbers come from
``numpy.random.default_rng(seed)``; grid edges are enumerated axis by axis
(all x-edges, then y-edges, then z-edges), each axis in increasing grid index
of the lower endpoint, with grid index ``x + sx*(y + sy*z)``.  Matrices are
returned as the lower triangle in scipy CSC form.
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["grid_edges", "spd_from_edges", "grid2d", "grid3d", "rgg", "band_arrow",
           "batch_grids", "two_cliques", "reference_test_matrix"]


def grid_edges(shape):
    """Edges (i, j) with i > j of the 2*d-point stencil on a grid of ``shape``
    (sx, sy[, sz]), enumerated axis by axis."""
    shape = tuple(int(s) for s in shape)
    dims = len(shape)
    idx = np.arange(int(np.prod(shape)), dtype=np.int64)
    coords = []
    rem = idx.copy()
    for s in shape:
        coords.append(rem % s)
        rem //= s
    stride = np.cumprod((1,) + shape[:-1])
    I, J = [], []
    for ax in range(dims):
        m = coords[ax] < shape[ax] - 1
        lo = idx[m]
        hi = lo + stride[ax]
        I.append(hi)
        J.append(lo)
    return np.concatenate(I), np.concatenate(J)


def spd_from_edges(n, I, J, rng):
    """Lower-triangular CSC matrix with a_ij = -U(0,1) on the edges and the
    diagonally dominant diagonal (row sum of |a_ij| plus 1)."""
    I = np.asarray(I, dtype=np.int64)
    J = np.asarray(J, dtype=np.int64)
    v = -rng.uniform(0.0, 1.0, size=len(I))
    d = np.ones(n)
    np.add.at(d, I, -v)
    np.add.at(d, J, -v)
    rows = np.concatenate([I, np.arange(n)])
    cols = np.concatenate([J, np.arange(n)])
    vals = np.concatenate([v, d])
    return sp.csc_matrix((vals, (rows, cols)), shape=(n, n))


def grid2d(side, seed=0, rng=None):
    """SPD matrix on a side x side 5-point grid (n = side**2)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    I, J = grid_edges((side, side))
    return spd_from_edges(side * side, I, J, rng)


def grid3d(side, seed=1, rng=None):
    """SPD matrix on a side^3 7-point grid (n = side**3)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    I, J = grid_edges((side, side, side))
    return spd_from_edges(side ** 3, I, J, rng)


def batch_grids(count=512, lo=16, hi=96, seed=2):
    """``count`` independent 2D 5-point grid matrices, side ~ U{lo..hi}; one
    generator stream draws the sides and then every instance's values."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=count)
    return [grid2d(int(s), rng=rng) for s in sides]


def rgg(n=200000, mean_degree=12.0, seed=5):
    """Random geometric graph: n points uniform in the unit cube joined to all
    neighbours within the radius that gives the requested mean degree (the
    radius is the length of the round(n*mean_degree/2)-th shortest pair)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, size=(n, 3))
    m = int(round(n * mean_degree / 2.0))
    r = (3.0 * mean_degree / (4.0 * np.pi * n)) ** (1.0 / 3.0) * 1.15
    tree = cKDTree(X)
    while True:
        pairs = tree.query_pairs(r, output_type="ndarray")
        if len(pairs) >= m:
            break
        r *= 1.1
    d = np.linalg.norm(X[pairs[:, 0]] - X[pairs[:, 1]], axis=1)
    keep = np.argsort(d, kind="stable")[:m]
    pairs = pairs[keep]
    pairs = pairs[np.lexsort((pairs.max(1), pairs.min(1)))]
    I = pairs.max(1)
    J = pairs.min(1)
    return spd_from_edges(n, I, J, rng)


def band_arrow(n=20000, bandwidth=8, arrows=50, seed=4):
    """Chordal band-plus-arrow pattern: a band of half-width ``bandwidth``
    plus ``arrows`` dense rows/columns at the end (vertices n-arrows..n-1)."""
    rng = np.random.default_rng(seed)
    nb = n - arrows
    I, J = [], []
    for d in range(1, bandwidth + 1):
        j = np.arange(nb - d, dtype=np.int64)
        I.append(j + d)
        J.append(j)
    for a in range(arrows):
        r = nb + a
        j = np.arange(r, dtype=np.int64)
        I.append(np.full(r, r, dtype=np.int64))
        J.append(j)
    return spd_from_edges(n, np.concatenate(I), np.concatenate(J), rng)


def two_cliques(n1=300, n2=100, n3=200, seed=6):
    """Nodes 0..n1-1 form a clique fully joined to the clique on
    n1..n1+n2+n3-1 through its first n2 nodes.  In natural order this gives a
    supernode of n1 columns with an n2-row separator and a root supernode of
    n2 + n3 columns: dense fronts larger than one 256-wide panel."""
    rng = np.random.default_rng(seed)
    def clique(lo, hi):
        idx = np.arange(lo, hi)
        I, J = np.meshgrid(idx, idx, indexing="ij")
        m = I > J
        return I[m], J[m]
    I1, J1 = clique(0, n1)
    I3, J3 = clique(n1, n1 + n2 + n3)
    a = np.arange(n1)
    b = np.arange(n1, n1 + n2)
    Ib, Jb = np.meshgrid(b, a, indexing="ij")
    I = np.concatenate([I1, Ib.ravel(), I3])
    J = np.concatenate([J1, Jb.ravel(), J3])
    return spd_from_edges(n1 + n2 + n3, I, J, rng)


def caterpillar(spine=400, bandwidth=6, arrows=10, seed=7):
    """A band of half-bandwidth `bandwidth` on `spine` vertices, each with a
    pendant leaf, plus `arrows` vertices joined to every spine vertex.  Leaves
    come first in the numbering (0..spine-1), then the spine, then the arrow
    vertices, so the natural order is a perfect elimination order: every
    spine supernode (one column, na = bandwidth + arrows) sits alone on its
    elimination-tree level and has two children, its leaf and the previous
    spine supernode."""
    rng = np.random.default_rng(seed)
    L, A = spine, arrows
    s = np.arange(L)
    I = [L + s]                    # spine vertex - its leaf
    J = [s]
    for d in range(1, bandwidth + 1):  # band along the spine
        I.append(L + s[d:])
        J.append(L + s[:-d])
    for a in range(A):             # arrow vertex - every spine vertex
        I.append(np.full(L, 2 * L + a))
        J.append(L + s)
    ia = np.arange(A)
    Ia, Ja = np.meshgrid(ia, ia, indexing="ij")
    m = Ia > Ja
    I.append(2 * L + Ia[m])        # the arrow vertices form a clique
    J.append(2 * L + Ja[m])
    return spd_from_edges(2 * L + A, np.concatenate(I), np.concatenate(J), rng)


def reference_test_matrix(n=31, density=0.15, seed=1):
    """The matrix of the reference's numeric test suite
    (tests/test_numeric.py:8-15 of the reference): ``random.sample`` of
    round(0.15 n^2) positions of an n x n grid, values (1..nnz)/nnz, lower
    triangle kept, plus 10 I."""
    import random
    nnz = int(round(density * n ** 2))
    random.seed(seed)
    nz = np.array(random.sample(range(n ** 2), nnz), dtype=np.int64)
    vals = np.arange(1, nnz + 1, dtype=np.float64) / nnz
    rows, cols = nz % n, nz // n
    m = rows >= cols
    A = sp.csc_matrix((vals[m], (rows[m], cols[m])), shape=(n, n))
    return A + sp.identity(n, format="csc") * 10.0
