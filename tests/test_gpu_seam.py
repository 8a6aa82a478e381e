"""GPU tests of the plugin seam's extra entry points and argument handling:
``frontal_add_update`` / ``frontal_get_update`` (cbase.c:11-59) against the
oracle's extend-add and extract, trsm / trmm with strided or host right-hand
sides (cbase.c:965-1041), and argument validation before any launch."""
import numpy as np
import pytest

import chompack_amd as cp
from chompack_amd import problems
from oracle import numeric as on

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _relidx_case():
    """relidx / relptr of a real symbolic factorization (the children of the
    largest front), with frontal matrices of the parent's clique order."""
    A = problems.grid2d(20, seed=0)
    s = cp.symbolic(A, p=cp.amd.order)
    nj = np.diff(s.sncolptr)
    k = int(np.argmax(nj))
    kids = [int(c) for c in s.chidx[s.chptr[k]:s.chptr[k + 1]]]
    return s, int(nj[k]), kids


@pytest.mark.parametrize("layout", ["rowmajor", "colmajor", "strided"])
def test_frontal_add_and_get_update(layout):
    import torch
    print("case frontal", layout, flush=True)
    s, nf, kids = _relidx_case()
    rng = np.random.default_rng(5)
    F0 = rng.standard_normal((nf, nf))
    if layout == "rowmajor":
        F = torch.as_tensor(F0.copy(), device="cuda:0")
    elif layout == "colmajor":
        F = torch.as_tensor(F0.copy(), device="cuda:0").t().contiguous().t()
    else:  # every second row and every third column of a larger array
        big = torch.zeros((2 * nf, 3 * nf), dtype=torch.float64, device="cuda:0")
        F = big[::2, 1::3]
        F.copy_(torch.as_tensor(F0))
    Fh = F0.copy()
    for c in kids:
        r = np.asarray(s.relidx[s.relptr[c]:s.relptr[c + 1]], dtype=np.int64)
        N = len(r)
        Uh = rng.standard_normal((N, N))
        U = torch.as_tensor(Uh, device="cuda:0")
        if layout != "rowmajor":
            U = U.t().contiguous().t()
        alpha = -0.625
        cp.frontal_add_update(F, U, s.relidx, s.relptr, c, alpha=alpha)
        on._extend_add(Fh, alpha * Uh, r)
        G = cp.frontal_get_update(F, s.relidx, s.relptr, c)
        torch.cuda.synchronize()
        assert G.shape == (N, N)
        # a gather: bit-exact against the oracle's extract of the GPU's own F
        # (F itself differs from the oracle's by fused multiply-add rounding)
        same = np.array_equal(G.cpu().numpy(), on._extract(F.cpu().numpy(), r))
        assert same
        assert _rel(G.cpu().numpy(), on._extract(Fh, r)) <= 1e-15
    e = _rel(F.cpu().numpy(), Fh)
    print("frontal_add_update rel err %.3e" % e, flush=True)
    assert e <= 1e-15  # one (possibly fused) multiply-add per entry and child
    if layout == "strided":
        # the entries of big outside the view are untouched
        mask = torch.ones_like(big, dtype=torch.bool)
        mask[::2, 1::3] = False
        assert float(big[mask].abs().max()) == 0.0


def test_frontal_argument_checks():
    import torch
    s, nf, kids = _relidx_case()
    c = kids[0]
    N = int(s.relptr[c + 1] - s.relptr[c])
    F = torch.zeros((nf, nf), dtype=torch.float64, device="cuda:0")
    with pytest.raises(ValueError):
        cp.frontal_add_update(F, torch.zeros((N + 1, N + 1), dtype=torch.float64,
                                             device="cuda:0"), s.relidx, s.relptr, c)
    small = torch.zeros((3, 3), dtype=torch.float64, device="cuda:0")
    with pytest.raises(IndexError):
        cp.frontal_get_update(small, s.relidx, s.relptr, c)
    with pytest.raises(IndexError):
        cp.frontal_get_update(F, s.relidx, s.relptr, len(s.relptr))
    bad = np.array([0, 2, 2], dtype=np.int64)
    with pytest.raises(ValueError):
        cp.frontal_get_update(F, bad, np.array([0, 3]), 0)


def _factor():
    import torch
    A = problems.grid2d(24, seed=0)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    cp.cholesky(X)
    torch.cuda.synchronize()
    return A, symb, X


def test_trsm_strided_and_host_right_hand_sides():
    import torch
    A, symb, L = _factor()
    n = symb.n
    rng = np.random.default_rng(9)
    Mh = rng.standard_normal((n, 4))
    # reference result: the column-major (n, 4) solve
    R = torch.as_tensor(np.asfortranarray(Mh)).to("cuda:0").t().contiguous().t()
    assert R.stride(0) == 1
    cp.trsm(L, R, alpha=0.5)
    cp.trsm(L, R, trans='T')
    want = R.cpu().numpy()
    # a column of a row-major matrix (stride 4): solved through a copy,
    # the neighbouring columns untouched
    M = torch.as_tensor(Mh.copy(), device="cuda:0")
    col = M[:, 2]
    cp.trsm(L, col, alpha=0.5)
    cp.trsm(L, col, trans='T')
    got = M.cpu().numpy()
    assert _rel(got[:, 2], want[:, 2]) <= 1e-13
    assert np.array_equal(got[:, [0, 1, 3]], Mh[:, [0, 1, 3]])
    # an (n, 1) slice of a row-major matrix
    M = torch.as_tensor(Mh.copy(), device="cuda:0")
    cp.trsm(L, M[:, 1:2], alpha=0.5)
    cp.trsm(L, M[:, 1:2], trans='T')
    assert _rel(M.cpu().numpy()[:, 1], want[:, 1]) <= 1e-13
    # host right-hand sides, updated in place: Fortran numpy, C numpy, and a
    # buffer-protocol view (as a cvxopt matrix is)
    for h in (np.asfortranarray(Mh.copy()), np.ascontiguousarray(Mh.copy())):
        cp.trsm(L, h, alpha=0.5)
        cp.trsm(L, h, trans='T')
        assert _rel(h, want) <= 1e-13
    base = np.asfortranarray(Mh.copy())
    mv = memoryview(base)
    cp.trsm(L, mv, alpha=0.5)
    cp.trsm(L, mv, trans='T')
    assert _rel(base, want) <= 1e-13
    # trmm undoes the solves on a host vector
    v = want[:, 0].copy()
    cp.trmm(L, v, trans='T')
    cp.trmm(L, v, alpha=2.0)
    assert _rel(v, Mh[:, 0]) <= 1e-12


def test_triangular_argument_checks():
    import torch
    A, symb, L = _factor()
    n = symb.n
    B = torch.zeros((n, 2), dtype=torch.float64, device="cuda:0").t().contiguous().t()
    with pytest.raises(ValueError):
        cp.trsm(L, B, offsetB=-1)
    with pytest.raises(ValueError):
        cp.trsm(L, B, nrhs=3)
    with pytest.raises(ValueError):
        cp.trsm(L, B, offsetB=n + 1)
    with pytest.raises(ValueError):
        cp.trsm(L, B, ldB=n - 1)
    with pytest.raises(TypeError):
        cp.trsm(L, np.zeros((n, 2), dtype=np.int64))


def test_pfcholesky_argument_checks():
    import torch
    from chompack_amd.pfcholesky import pfchol, pftrsm
    n, m = 50, 4
    dev = "cuda:0"
    cm = lambda r, c: torch.zeros((c, r), dtype=torch.float64, device=dev).t()
    V = cm(n, m)
    with pytest.raises(ValueError):
        pfchol(np.ones(m), V, cm(n, m - 1), cm(n, m))
    with pytest.raises(ValueError):
        pfchol(np.ones(m), V, cm(n, m), cm(n - 1, m))
    with pytest.raises(ValueError):
        pftrsm(V, cm(n, m), cm(n, m), cm(n + 1, 2))
