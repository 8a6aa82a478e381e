"""GPU parity: HIP sweeps vs the NumPy oracle on the same symbolic object.

Tolerance (FP64 path): max |gpu - oracle| / max |oracle| <= 1e-12 over the
whole blkval array (reordered sums in the dense kernels move the last bits).
Each problem runs twice: with the default LDS/large-front split and with
``small_max = 1``, which routes every front through the large-front kernels
(column-tiled extend-add, panel factorization, MFMA GEMM tiles).
"""
import numpy as np
import pytest

import chompack_amd as cp
from chompack_amd import numeric, problems
from oracle import numeric as on

pytestmark = pytest.mark.gpu

TOL = 1e-12

PROBLEMS = {
    "ref31": lambda: problems.reference_test_matrix(),
    "grid2d_64": lambda: problems.grid2d(64, seed=0),
    "grid3d_12": lambda: problems.grid3d(12, seed=1),
    "two_cliques": lambda: problems.two_cliques(),
    # an elimination-tree chain of one-column fronts with two children each
    # (the completion's chain kernel with non-chain children)
    "caterpillar": lambda: problems.caterpillar(),
    # an elimination-tree path of one-column fronts (natural order): the
    # chain launches of cholesky, llt and the projected inverse
    "band_arrow": lambda: problems.band_arrow(1500, 8, 12, seed=4),
}
ORDER = {"two_cliques": None, "caterpillar": None, "band_arrow": None}


def _setup(name, small_max):
    import torch
    A = PROBLEMS[name]()
    symb = cp.symbolic(A, p=ORDER.get(name, cp.amd.order))
    symb._plans["cuda:0"] = numeric.Plan(symb, "cuda:0", small_max=small_max)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    torch.cuda.synchronize()
    return A, symb, X


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("small_max", [64, 1])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_cholesky_projinv_solve(name, small_max):
    import torch
    print("case", name, small_max, flush=True)
    A, symb, X = _setup(name, small_max)
    ref = X.blkval.cpu().numpy().copy()
    A0 = ref.copy()
    assert on.cholesky(symb, ref) == 0
    cp.cholesky(X)
    assert X.is_factor is True
    e = _rel(X.blkval.cpu().numpy(), ref)
    print("cholesky rel err %.3e" % e, flush=True)
    assert e <= TOL

    # solve with the factor: trsm N then T equals A^{-1} b
    rng = np.random.default_rng(7)
    Bh = rng.standard_normal((symb.n, 3))
    B = torch.as_tensor(Bh.copy(), device="cuda:0")
    Bref = np.asfortranarray(Bh.copy())
    on.trsm(symb, ref, Bref, trans="N")
    cp.trsm(X, B)
    e = _rel(B.cpu().numpy(), Bref)
    print("trsm N rel err %.3e" % e, flush=True)
    assert e <= TOL
    on.trsm(symb, ref, Bref, trans="T")
    cp.trsm(X, B, trans="T")
    e = _rel(B.cpu().numpy(), Bref)
    print("trsm T rel err %.3e" % e, flush=True)
    assert e <= TOL
    As = A + A.T - np.diag(A.diagonal())
    res = np.abs(As @ B.cpu().numpy() - Bh).max() / np.abs(Bh).max()
    assert res <= 1e-12

    # products with the factor (trmm N / T, alpha) against the oracle; 20
    # right-hand sides (a 16-wide chunk and a partial one in the small-front
    # kernels)
    Mh = rng.standard_normal((symb.n, 20))
    for tr in "NT":
        M = torch.as_tensor(np.asfortranarray(Mh)).to("cuda:0").t().contiguous().t()
        Mref = np.asfortranarray(Mh.copy())
        on.trmm(symb, ref, Mref, alpha=0.75, trans=tr)
        cp.trmm(X, M, alpha=0.75, trans=tr)
        e = _rel(M.cpu().numpy(), Mref)
        print("trmm %s rel err %.3e" % (tr, e), flush=True)
        assert e <= TOL
        cp.trsm(X, M, alpha=1.0 / 0.75, trans=tr)   # undone by the solve
        assert _rel(M.cpu().numpy(), Mh) <= 1e-12

    on.projected_inverse(symb, ref)
    cp.projected_inverse(X)
    assert X.is_factor is False
    e = _rel(X.blkval.cpu().numpy(), ref)
    print("projected_inverse rel err %.3e" % e, flush=True)
    assert e <= TOL
    # <P(A^-1), A> = n  (tests/test_numeric.py:40-46 of the reference)
    Y = X
    Xa = cp.cspmatrix(symb, device="cuda:0") + A
    v = cp.dot(Y, Xa)
    assert abs(v - symb.n) <= 1e-9 * symb.n
    assert abs(v - on.dot(symb, Y.blkval.cpu().numpy(), A0)) <= 1e-10 * symb.n


@pytest.mark.parametrize("small_max", [64, 1])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_llt_roundtrip(name, small_max):
    print("case llt", name, small_max, flush=True)
    A, symb, X = _setup(name, small_max)
    A0 = X.blkval.cpu().numpy().copy()
    ref = A0.copy()
    on.cholesky(symb, ref)
    on.llt(symb, ref)
    cp.cholesky(X)
    cp.llt(X)
    assert X.is_factor is False
    e = _rel(X.blkval.cpu().numpy(), ref)
    assert e <= TOL
    assert _rel(X.blkval.cpu().numpy(), A0) <= 1e-12


def test_cholesky_failure_raises():
    import scipy.sparse as sp
    A = problems.grid2d(8, seed=0)
    A = sp.csc_matrix(A - 30.0 * sp.identity(64))
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    with pytest.raises(ArithmeticError, match="factorization failed"):
        cp.cholesky(X)
    assert X.is_factor is True


@pytest.mark.parametrize("factored", [True, False])
def test_completion_failure_raises(factored):
    """A partial matrix with a negative diagonal entry in the middle of the
    caterpillar's chain is not completable: the reversed Cholesky of that
    front's (1,1) block meets a non-positive pivot.  The oracle fails as
    well; the GPU sweep (the chain kernel with factored updates, the level
    sweep without) reports it as ArithmeticError."""
    import scipy.sparse as sp
    A = problems.caterpillar()
    d = np.zeros(A.shape[0])
    d[400 + 200] = -A.diagonal()[400 + 200] - 10.0  # spine vertex 200
    A = sp.csc_matrix(A + sp.diags(d))
    symb = cp.symbolic(A, p=None)
    ref = (cp.cspmatrix(symb, device="cpu") + A).blkval.numpy().copy()
    assert on.completion(symb, ref, factored) != 0
    X = cp.cspmatrix(symb, device="cuda:0") + A
    with pytest.raises(ArithmeticError, match="completion failed"):
        cp.completion(X, factored_updates=factored)
    assert X.is_factor is True


def test_graph_replays_match_eager():
    """The sweeps replay a captured HIP graph; every replay (with eager work
    and synchronizations in between) must equal the eager launch sequence."""
    import torch
    from chompack_amd import _native
    A = problems.grid3d(10, seed=4)
    symb = cp.symbolic(A, p=cp.amd.order)
    symb._plans["cuda:0"] = numeric.Plan(symb, "cuda:0", small_max=32)
    plan = symb._plans["cuda:0"]
    X0 = cp.cspmatrix(symb, device="cuda:0") + A
    lib = _native.lib()
    lib.chp_set_graphs(plan.handle, 0)
    E = X0.copy()
    cp.cholesky(E)
    cp.projected_inverse(E)
    eager = E.blkval.cpu().numpy()
    lib.chp_set_graphs(plan.handle, 1)
    X = cp.cspmatrix(symb, blkval=torch.empty_like(X0.blkval))
    for rep in range(4):
        X.blkval.copy_(X0.blkval)
        X.is_factor = False
        cp.cholesky(X)
        cp.projected_inverse(X)
        torch.cuda.synchronize()
        assert np.array_equal(X.blkval.cpu().numpy(), eager), rep
        _ = torch.ones(1000, device="cuda:0").sum().item()  # eager work between replays


def _lower_mask(symb):
    m = np.zeros(int(symb.blkptr[-1]), dtype=bool)
    for k in range(symb.Nsn):
        nn = int(symb.snptr[k + 1] - symb.snptr[k])
        nj = nn + int(symb.relptr[k + 1] - symb.relptr[k])
        blk = m[symb.blkptr[k]:symb.blkptr[k + 1]].reshape((nn, nj)).T
        blk[np.tril_indices(nj, 0, nn)] = True
    return m


def _factor_pair(name):
    """Oracle factor L and projected inverse Y of a test problem."""
    A = PROBLEMS[name]()
    symb = cp.symbolic(A, p=ORDER.get(name, cp.amd.order))
    b = (cp.cspmatrix(symb, device="cpu") + A).blkval.numpy().copy()
    L = b.copy()
    assert on.cholesky(symb, L) == 0
    Y = L.copy()
    assert on.projected_inverse(symb, Y) == 0
    return symb, L, Y


@pytest.mark.parametrize("factored", [True, False])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_completion(name, factored):
    import torch
    print("case completion", name, factored, flush=True)
    symb, L, Y = _factor_pair(name)
    X = cp.cspmatrix(symb, blkval=torch.as_tensor(Y.copy(), device="cuda:0"))
    cp.completion(X, factored_updates=factored)
    assert X.is_factor is True
    ref = Y.copy()
    assert on.completion(symb, ref, factored) == 0
    m = _lower_mask(symb)
    got = X.blkval.cpu().numpy()
    e = _rel(got[m], ref[m])
    print("completion rel err %.3e" % e, flush=True)
    assert e <= TOL
    assert _rel(got[m], L[m]) <= 1e-11   # completion(P(X^-1)) recovers the factor


@pytest.mark.parametrize("factored", [True, False])
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_hessian_modes(name, factored):
    import torch
    symb, L, Y = _factor_pair(name)
    Lc = cp.cspmatrix(symb, blkval=torch.as_tensor(L, device="cuda:0"), factor=True)
    Yc = cp.cspmatrix(symb, blkval=torch.as_tensor(Y, device="cuda:0"))
    rng = np.random.default_rng(11)
    for adj in (False, True, None):
        for inv in (False, True):
            print("case hessian", name, factored, adj, inv, flush=True)
            us = [rng.standard_normal(len(L)) for _ in range(2)]
            Us = [cp.cspmatrix(symb, blkval=torch.as_tensor(u.copy(), device="cuda:0")) for u in us]
            cp.hessian(Lc, Yc, Us, adj=adj, inv=inv, factored_updates=factored)
            refs = [u.copy() for u in us]
            on.hessian(symb, L, Y, refs, adj=adj, inv=inv, factored_updates=factored)
            m = _lower_mask(symb)
            for U, r in zip(Us, refs):
                e = _rel(U.blkval.cpu().numpy()[m], r[m])
                print("hessian rel err %.3e" % e, flush=True)
                assert e <= TOL, (adj, inv)


@pytest.mark.parametrize("factored", [True, False])
def test_hessian_direction_chunks(factored):
    """35 directions: the Hessian maps run on chunks of directions stacked on
    the device (32, then 3); every direction against the oracle, for the
    full map (adj=None) and its inverse, and G_X^adj."""
    import torch
    symb, L, Y = _factor_pair("grid3d_12")
    Lc = cp.cspmatrix(symb, blkval=torch.as_tensor(L, device="cuda:0"), factor=True)
    Yc = cp.cspmatrix(symb, blkval=torch.as_tensor(Y, device="cuda:0"))
    rng = np.random.default_rng(12)
    m = _lower_mask(symb)
    for adj, inv in ((None, False), (None, True), (True, False)):
        print("case hessian chunks", factored, adj, inv, flush=True)
        us = [rng.standard_normal(len(L)) for _ in range(35)]
        Us = [cp.cspmatrix(symb, blkval=torch.as_tensor(u.copy(), device="cuda:0")) for u in us]
        cp.hessian(Lc, Yc, Us, adj=adj, inv=inv, factored_updates=factored)
        refs = [u.copy() for u in us]
        on.hessian(symb, L, Y, refs, adj=adj, inv=inv, factored_updates=factored)
        e = max(_rel(U.blkval.cpu().numpy()[m], r[m]) for U, r in zip(Us, refs))
        print("hessian chunks max rel err %.3e" % e, flush=True)
        assert e <= TOL, (adj, inv)


def test_logdet_and_gradient():
    """logdet from the GPU factor vs the dense slogdet; the autograd gradient
    (projected inverse) vs central finite differences, relative error <= 1e-6."""
    import torch
    A = problems.grid3d(8, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    L = X.copy()
    cp.cholesky(L)
    Ad = cp.symmetrize(A).toarray()
    sign, ref = np.linalg.slogdet(Ad)
    assert sign > 0 and abs(cp.logdet(L) - ref) <= 1e-12 * abs(ref)
    m = torch.as_tensor(_lower_mask(symb), device="cuda:0")
    rng = np.random.default_rng(4)
    U = torch.as_tensor(rng.standard_normal(len(m)), device="cuda:0") * m
    Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
    val = cp.autograd.logdet(Xg)
    assert abs(val.item() - ref) <= 1e-12 * abs(ref)
    val.backward()
    dd = float((Xg.blkval.grad * U).sum())
    h = 1e-5
    f = lambda t: cp.autograd.logdet(cp.cspmatrix(symb, blkval=X.blkval + t * U)).item()
    fd = (f(h) - f(-h)) / (2 * h)
    print("directional derivative %.12e fd %.12e" % (dd, fd), flush=True)
    assert abs(dd - fd) <= 1e-6 * abs(dd)


def test_grid40_full_size_vs_reference():
    """The 40^3 7-point grid (n = 64000, seed 1, AMD) at full size: cholesky,
    the solve with 16 N(0,1) right-hand sides (L^{-1} then L^{-T}) and the
    projected inverse against the reference's own C kernels (oracle/_ref),
    relative error <= 1e-12; plus the size-independent identities
    ||A x - b|| / ||b|| and <P(A^{-1}), A> = n."""
    import torch
    from oracle import ref
    A = problems.grid3d(40, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    b0 = X.blkval.cpu().numpy().copy()
    cp.cholesky(X)
    rng = np.random.default_rng(0)
    Bh = rng.standard_normal((symb.n, 16))
    B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()
    cp.trsm(X, B)
    cp.trsm(X, B, trans="T")
    xs = B.cpu().numpy()
    As = cp.symmetrize(A).tocsr()
    res = np.linalg.norm(As @ xs - Bh, axis=0) / np.linalg.norm(Bh, axis=0)
    print("grid40 solve residual %.3e" % res.max(), flush=True)
    assert res.max() <= 1e-12
    Lg = X.blkval.cpu().numpy().copy()
    cp.projected_inverse(X)
    Yg = X.blkval.cpu().numpy()
    Xa = cp.cspmatrix(symb, device="cuda:0") + A
    assert abs(cp.dot(X, Xa) - symb.n) <= 1e-9 * symb.n
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    m = _lower_mask(symb)
    L = b0.copy()
    assert ref.cholesky(symb, L) == 0
    e = _rel(Lg[m], L[m])
    print("grid40 cholesky rel err vs reference %.3e" % e, flush=True)
    assert e <= TOL
    Br = np.asfortranarray(Bh.copy())
    ref.trsm(symb, L, Br, trans="N")
    ref.trsm(symb, L, Br, trans="T")
    e = _rel(xs, Br)
    print("grid40 solve rel err vs reference %.3e" % e, flush=True)
    assert e <= TOL
    assert ref.projected_inverse(symb, L) == 0
    e = _rel(Yg[m], L[m])
    print("grid40 projected_inverse rel err vs reference %.3e" % e, flush=True)
    assert e <= TOL


def test_solve_graph_replays_match_eager():
    """Repeated solves with the same buffers replay a captured HIP graph; each
    replay must equal the eager solve, also after the work buffers grow."""
    import torch
    A = problems.grid3d(10, seed=6)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    cp.cholesky(X)
    rng = np.random.default_rng(3)
    B0 = torch.as_tensor(np.asfortranarray(rng.standard_normal((symb.n, 4)))).to("cuda:0")
    B0 = B0.t().contiguous().t()
    outs = []
    B = B0.clone()
    for rep in range(5):
        B.copy_(B0)
        cp.trsm(X, B)
        cp.trsm(X, B, trans="T")
        outs.append(B.cpu().numpy().copy())
        _ = torch.ones(100, device="cuda:0").sum().item()
        if rep == 2:  # a wider solve grows the work buffer: graphs must be rebuilt
            W = torch.as_tensor(rng.standard_normal((symb.n, 64))).to("cuda:0").t().contiguous().t()
            cp.trsm(X, W)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_solve_after_refactor_in_same_buffer():
    """The solve keeps L11^{-1} of its large fronts between calls; a new
    factor written into the same blkval (a refactorization with other values,
    and a copy of a third factor) must give that factor's solution, with 20
    right-hand sides (a 16-wide chunk and a partial one)."""
    import torch
    A1 = problems.grid3d(12, seed=6)
    symb = cp.symbolic(A1, p=cp.amd.order)
    A2 = A1.copy()
    A2.data = A2.data * 1.5
    A2.setdiag(A2.diagonal() + 1.0)
    As = [cp.symmetrize(A).tocsr() for A in (A1, A2)]
    X = cp.cspmatrix(symb, device="cuda:0") + A1
    X2 = cp.cspmatrix(symb, device="cuda:0") + A2
    cp.cholesky(X2)
    rng = np.random.default_rng(5)
    Bh = rng.standard_normal((symb.n, 20))
    B0 = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()

    def solve_and_check(which):
        B = B0.clone().t().contiguous().t()
        cp.trsm(X, B)
        cp.trsm(X, B, trans="T")
        x = B.cpu().numpy()
        r = np.linalg.norm(As[which] @ x - Bh, axis=0) / np.linalg.norm(Bh, axis=0)
        assert r.max() <= 1e-12, (which, r.max())

    cp.cholesky(X)
    solve_and_check(0)
    solve_and_check(0)
    X.blkval.zero_()
    X.is_factor = False
    X += A2
    cp.cholesky(X)
    solve_and_check(1)
    X.blkval.copy_(X2.blkval)  # values changed behind the library's back
    solve_and_check(1)
    X.blkval.zero_()
    X.is_factor = False
    X += A1
    cp.cholesky(X)
    solve_and_check(0)


def test_batched_512_grids_logdet_and_solve():
    """512 independent 2D grids, side ~ U{16..96} (seed 2), as one batched
    factorization: per-instance logdet against the dense slogdet (the 24
    smallest instances) and against single-instance factorizations; the joint
    solve's residual per instance."""
    import torch
    As = problems.batch_grids(512, 16, 96, seed=2)
    bs = cp.batch.BatchSymbolic(As)
    X = bs.assemble()
    cp.cholesky(X)
    ld = bs.logdet(X)
    order = np.argsort([A.shape[0] for A in As])
    for i in order[:24]:
        sign, ref = np.linalg.slogdet(cp.symmetrize(As[i]).toarray())
        assert sign > 0 and abs(ld[i] - ref) <= 1e-12 * abs(ref), i
    for i in order[-3:]:
        s = cp.symbolic(As[i], p=cp.amd.order)
        Xi = cp.cspmatrix(s, device="cuda:0") + As[i]
        cp.cholesky(Xi)
        li = cp.logdet(Xi)
        assert abs(li - ld[i]) <= 1e-12 * abs(li), i
    rng = np.random.default_rng(8)
    bh = rng.standard_normal(bs.symb.n)
    b = torch.as_tensor(bh, device="cuda:0")
    cp.trsm(X, b)
    cp.trsm(X, b, trans="T")
    x = b.cpu().numpy()
    for A, xi, bi in zip(As, bs.split(x), bs.split(bh)):
        r = np.linalg.norm(cp.symmetrize(A) @ xi - bi) / np.linalg.norm(bi)
        assert r <= 1e-12
    # the same batch with FP32 Schur updates: logdet error against FP64
    Xm = bs.assemble()
    cp.mixed.cholesky(Xm)
    ldm = bs.logdet(Xm)
    err = np.abs(ldm - ld) / np.abs(ld)
    print("batched mixed logdet rel err max %.3e" % err.max(), flush=True)
    assert err.max() <= 1e-6


def test_grid32_gradient_and_hessian_directions():
    """-logdet on the 32^3 7-point grid (n = 32768, seed 3): the autograd
    gradient (projected-inverse backward) against central differences with
    step 1e-6 (relative error <= 1e-6) and against the reference's projected
    inverse; the full Hessian map on 32 random directions: symmetric under
    the trace inner product, equal to the derivative of the projected inverse
    (central differences) and to the reference's hessian()."""
    import time
    import torch
    from oracle import ref
    A = problems.grid3d(32, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    m = torch.as_tensor(_lower_mask(symb), device="cuda:0")
    w = torch.where(m, 1.0, 0.0).double()
    rng = np.random.default_rng(9)
    Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
    f = -cp.autograd.logdet(Xg)
    f.backward()
    G = Xg.blkval.grad
    U = torch.as_tensor(rng.standard_normal(len(m)), device="cuda:0") * w
    dd = float((G * U).sum())
    h = 1e-6

    def logdiag_at(t):
        Lt = cp.cspmatrix(symb, blkval=X.blkval + t * U)
        cp.cholesky(Lt)
        return cp.numeric.logdiag(Lt)
    # f(X + hU) - f(X - hU) summed column by column: the per-column log
    # differences are O(h), so the central difference keeps its digits
    # (the totals are ~1e5 and would cancel to ~1e-5 relative accuracy)
    fd = -2.0 * float((logdiag_at(h) - logdiag_at(-h)).sum()) / (2 * h)
    print("grid32 directional derivative %.10e fd %.10e" % (dd, fd), flush=True)
    assert abs(dd - fd) <= 1e-6 * abs(dd)
    # factor and projected inverse for the Hessian
    L = X.copy()
    cp.cholesky(L)
    Y = L.copy()
    cp.projected_inverse(Y)
    # gradient of -logdet is -Y mapped to blkval coordinates
    wgt = cp.autograd._grad_weights(symb, G.device)
    assert float((G + Y.blkval * wgt).abs().max()) <= 1e-12 * float(Y.blkval.abs().max())
    Us = [cp.cspmatrix(symb, blkval=torch.as_tensor(rng.standard_normal(len(m)), device="cuda:0") * w)
          for _ in range(32)]
    U0 = [u.blkval.clone() for u in Us]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cp.hessian(L, Y, Us, adj=None)
    torch.cuda.synchronize()
    print("grid32 hessian on 32 directions: %.1f ms" % ((time.perf_counter() - t0) * 1e3), flush=True)
    Hs = [u.blkval for u in Us]
    d = lambda a, b: cp.dot(cp.cspmatrix(symb, blkval=a), cp.cspmatrix(symb, blkval=b))
    for i, j in [(0, 1), (2, 7), (30, 31)]:
        a, b = d(Hs[i], U0[j]), d(U0[i], Hs[j])
        assert abs(a - b) <= 1e-11 * max(abs(a), 1.0)
    for i in (0, 5):
        def Yat(t):
            Xt = cp.cspmatrix(symb, blkval=X.blkval + t * U0[i])
            cp.cholesky(Xt)
            cp.projected_inverse(Xt)
            return Xt.blkval
        dY = (Yat(h) - Yat(-h)) / (2 * h)
        e = float(((dY + Hs[i]) * w).abs().max() / (Hs[i] * w).abs().max())
        print("grid32 hessian vs FD of Y: %.3e" % e, flush=True)
        assert e <= 1e-6
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    b0 = X.blkval.cpu().numpy().copy()
    Lr = b0.copy()
    assert ref.cholesky(symb, Lr) == 0
    Yr = Lr.copy()
    assert ref.projected_inverse(symb, Yr) == 0
    mm = _lower_mask(symb)
    assert _rel(Y.blkval.cpu().numpy()[mm], Yr[mm]) <= TOL
    urs = [u.cpu().numpy().copy() for u in U0[:4]]
    ref.hessian(symb, Lr, Yr, urs, adj=None)
    for a, r in zip(Hs[:4], urs):
        e = _rel(a.cpu().numpy()[mm], r[mm])
        print("grid32 hessian rel err vs reference %.3e" % e, flush=True)
        assert e <= TOL


def test_band_arrow_20000_completion():
    """Maximum-determinant completion on the chordal band-plus-arrow pattern
    (n = 20000, half-bandwidth 8, 50 arrow columns, seed 4, natural order: an
    elimination-tree chain of 19942 supernodes): completion(P(A^{-1})) is the
    factor of A, and equals the reference's completion."""
    import time
    import torch
    from oracle import ref
    A = problems.band_arrow(20000, 8, 50, seed=4)
    symb = cp.symbolic(A)
    assert symb.fill == (0, 0)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    cp.cholesky(X)
    Lg = X.blkval.cpu().numpy().copy()
    cp.projected_inverse(X)
    Yg = X.blkval.cpu().numpy().copy()
    for factored in (True, False):
        Z = cp.cspmatrix(symb, blkval=torch.as_tensor(Yg, device="cuda:0"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cp.completion(Z, factored_updates=factored)
        torch.cuda.synchronize()
        print("band_arrow completion factored=%s: %.1f ms" % (factored, (time.perf_counter() - t0) * 1e3),
              flush=True)
        mm = _lower_mask(symb)
        got = Z.blkval.cpu().numpy()
        assert _rel(got[mm], Lg[mm]) <= 1e-11
        if ref.available():
            r = Yg.copy()
            assert ref.completion(symb, r, factored) == 0
            assert _rel(got[mm], r[mm]) <= TOL


def test_rgg_200k_cholesky_and_64rhs_solve():
    """Random geometric graph, n = 200000 points in the unit cube, mean degree
    12 (seed 5), AMD: cholesky against the reference's kernels, and the solve
    with 64 right-hand sides: residual ||A x - b|| / ||b|| per column."""
    import torch
    from oracle import ref
    A = problems.rgg(200000, 12.0, seed=5)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    b0 = X.blkval.cpu().numpy()
    cp.cholesky(X)
    rng = np.random.default_rng(0)
    Bh = rng.standard_normal((symb.n, 64))
    B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()
    cp.trsm(X, B)
    cp.trsm(X, B, trans="T")
    As = cp.symmetrize(A).tocsr()
    res = np.linalg.norm(As @ B.cpu().numpy() - Bh, axis=0) / np.linalg.norm(Bh, axis=0)
    print("rgg residual %.3e" % res.max(), flush=True)
    assert res.max() <= 1e-12
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    L = b0.copy()
    assert ref.cholesky(symb, L) == 0
    m = _lower_mask(symb)
    e = _rel(X.blkval.cpu().numpy()[m], L[m])
    print("rgg cholesky rel err vs reference %.3e" % e, flush=True)
    assert e <= TOL


def test_mixed_precision_refined_solve():
    """FP32 Schur updates + FP64 iterative refinement: the refined solve
    reaches max_j ||b_j - A x_j|| / ||b_j|| <= 1e-10; the mixed factor's
    logdet is within 1e-6 relative of the FP64 one, its factor within 1e-5
    relative (FP32 rounding of the updates, each element accumulated in FP32)."""
    import torch
    A = problems.grid3d(16, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    symb._plans["cuda:0"] = numeric.Plan(symb, "cuda:0", small_max=16)  # exercise large fronts
    X0 = cp.cspmatrix(symb, device="cuda:0") + A
    L64 = X0.copy()
    cp.cholesky(L64)
    L32 = X0.copy()
    cp.mixed.cholesky(L32)
    assert L32.is_factor is True
    e = float((L32.blkval - L64.blkval).abs().max() / L64.blkval.abs().max())
    print("mixed factor rel diff %.3e" % e, flush=True)
    assert 0.0 < e <= 1e-5   # FP32 arithmetic was used, and only at FP32 level
    l64, l32 = cp.logdet(L64), cp.logdet(L32)
    assert abs(l32 - l64) <= 1e-6 * abs(l64)
    Am = cp.mixed.SpMat(A)
    rng = np.random.default_rng(2)
    Bh = rng.standard_normal((symb.n, 4))
    B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0")
    x, hist = cp.mixed.solve(Am, L32, B, tol=1e-10)
    print("refinement", hist, flush=True)
    assert hist[-1] <= 1e-10
    assert np.abs(B.cpu().numpy() - Bh).max() == 0.0   # B untouched
    As = cp.symmetrize(A).tocsr()
    res = np.linalg.norm(As @ x.cpu().numpy() - Bh, axis=0) / np.linalg.norm(Bh, axis=0)
    assert res.max() <= 1e-10


MIXED_CASES = {
    # 40^3 grid (seed 1), 16 right-hand sides; RGG n = 200000 (seed 5), 64
    "grid40_16rhs": (lambda: problems.grid3d(40, seed=1), 16),
    "rgg200k_64rhs": (lambda: problems.rgg(200000, 12.0, seed=5), 64),
}


@pytest.mark.parametrize("schur", ["fp32", "bf16x3"])
@pytest.mark.parametrize("case", list(MIXED_CASES))
def test_mixed_refined_solve_full_size(case, schur):
    """The mixed-precision legs at full size: cholesky with FP32 Schur
    updates, then FP64 iterative refinement of N(0,1) right-hand sides to
    max_j ||b_j - A x_j||_2 / ||b_j||_2 <= 1e-10 (checked again on the host
    with scipy).  The mixed factor's difference from the FP64 factor is
    stated and bounded by 1e-4 relative for FP32 Schur updates (rounding of
    long Schur sums) and 1e-2 for BF16x3 (about 2^-16 per product); its
    logdet by 1e-7 and 1e-5 relative."""
    import torch
    make, nrhs = MIXED_CASES[case]
    print("case mixed", case, schur, flush=True)
    A = make()
    symb = cp.symbolic(A, p=cp.amd.order)
    X0 = cp.cspmatrix(symb, device="cuda:0") + A
    L64 = X0.copy()
    cp.cholesky(L64)
    L32 = X0.copy()
    cp.mixed.cholesky(L32, schur=schur)
    del X0
    e = float((L32.blkval - L64.blkval).abs().max() / L64.blkval.abs().max())
    l64, l32 = cp.logdet(L64), cp.logdet(L32)
    print("mixed factor rel diff %.3e, logdet rel diff %.3e" % (e, abs(l32 - l64) / abs(l64)),
          flush=True)
    assert 0.0 < e <= (1e-4 if schur == "fp32" else 1e-2)
    assert abs(l32 - l64) <= (1e-7 if schur == "fp32" else 1e-5) * abs(l64)
    del L64
    Am = cp.mixed.SpMat(A)
    rng = np.random.default_rng(0)
    Bh = rng.standard_normal((symb.n, nrhs))
    B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()
    x, hist = cp.mixed.solve(Am, L32, B, tol=1e-10)
    print("refinement history", ["%.2e" % h for h in hist], flush=True)
    assert hist[-1] <= 1e-10
    assert len(hist) <= (4 if schur == "fp32" else 8)
    As = cp.symmetrize(A).tocsr()
    res = np.linalg.norm(As @ x.cpu().numpy() - Bh, axis=0) / np.linalg.norm(Bh, axis=0)
    print("host residual %.3e" % res.max(), flush=True)
    assert res.max() <= 1e-10


@pytest.mark.parametrize("nrhs", [1, 5, 70])
def test_spmv_residual(nrhs):
    """R = B - A X on the GPU (device CSR, row-major copy of X for several
    right-hand sides, 64-column chunks) against scipy, on the random
    geometric graph's pattern (no locality in the column indices)."""
    import torch
    A = problems.rgg(3000, 12.0, seed=5)
    Am = cp.mixed.SpMat(A)
    As = cp.symmetrize(A).tocsr()
    rng = np.random.default_rng(nrhs)
    Xh = rng.standard_normal((A.shape[0], nrhs))
    Bh = rng.standard_normal((A.shape[0], nrhs))
    col = lambda M: torch.as_tensor(np.asfortranarray(M)).to("cuda:0").t().contiguous().t()
    X, B = col(Xh), col(Bh)
    if nrhs == 1:
        X, B = X[:, 0].contiguous(), B[:, 0].contiguous()
    R = torch.empty_like(X)
    Am.residual(X, B, R)
    ref = (Bh - As @ Xh).reshape(R.shape)
    err = np.abs(R.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-14, err


def test_pfcholesky():
    """Product-form Cholesky: the scan kernels against the oracle's
    sequential restatement of pfcholesky.c (factor, solve, product, both
    transposes), the factored operator against X + V diag(a) V^T (solve
    residual, product round trip), and the reference's failure on an
    indefinite downdate."""
    import torch
    from chompack_amd.pfcholesky import pfcholesky, pfchol, pftrsm, pftrmm
    rng = np.random.default_rng(12)
    n, m = 2500, 6
    V0 = rng.standard_normal((n, m))
    a = rng.uniform(0.5, 2.0, m)
    a[1] = -1e-4   # a downdate that stays positive definite (1 / 1e-4 > |v|^2)
    col = lambda M: torch.as_tensor(np.asfortranarray(M)).to("cuda:0").t().contiguous().t()
    V, L, B = col(V0), col(np.zeros((n, m))), col(np.zeros((n, m)))
    pfchol(a, V, L, B)
    Vr, Lr, Br = V0.copy(), np.zeros((n, m)), np.zeros((n, m))
    assert on.pf_factor(a, Vr, Lr, Br) == 0
    for got, want in ((V, Vr), (L, Lr), (B, Br)):
        assert _rel(got.cpu().numpy(), want) <= TOL
    U0 = rng.standard_normal((n, 5))
    for tr in "NT":
        for gfun, ofun in ((pftrsm, on.pf_solve), (pftrmm, on.pf_product)):
            U = col(U0)
            gfun(V, L, B, U, trans=tr)
            Ur = U0.copy()
            ofun(Vr, Lr, Br, Ur, trans=tr)
            assert _rel(U.cpu().numpy(), Ur) <= TOL, (gfun.__name__, tr)
    # the whole factorization of X + V diag(a) V^T on a chordal X
    A = problems.grid2d(50, seed=2)
    W0 = rng.standard_normal((A.shape[0], 4))
    aw = np.array([1.0, 0.5, 2.0, 0.25])
    F = pfcholesky(A, W0, a=aw, p=cp.amd.order)
    M = cp.symmetrize(A).toarray() + W0 @ np.diag(aw) @ W0.T
    bh = rng.standard_normal((A.shape[0], 3))
    x = col(bh)
    F.trsm(x)
    F.trsm(x, trans="T")
    r = np.linalg.norm(M @ x.cpu().numpy() - bh, axis=0) / np.linalg.norm(bh, axis=0)
    assert r.max() <= 1e-12
    y = col(bh)
    F.trmm(y, trans="T")
    F.trmm(y)
    assert _rel(y.cpu().numpy(), M @ bh) <= 1e-12
    with pytest.raises(ArithmeticError, match="factorization failed"):
        pfcholesky(A, W0, a=-10.0 * np.ones(4), p=cp.amd.order)


def _edge_problems():
    import scipy.sparse as sp
    rng = np.random.default_rng(21)
    out = {}
    out["n1"] = sp.csc_matrix(np.array([[3.0]]))
    out["diagonal"] = sp.csc_matrix(sp.diags(rng.uniform(1.0, 2.0, 100)))
    D = rng.standard_normal((200, 200))
    out["dense200"] = sp.csc_matrix(np.tril(D @ D.T + 200.0 * np.eye(200)))
    n = 300   # tridiagonal: an elimination-tree chain of singleton fronts
    T = sp.diags([np.full(n - 1, -1.0), np.full(n, 3.0)], [-1, 0], format="csc")
    out["chain"] = T
    return out


@pytest.mark.parametrize("name", ["n1", "diagonal", "dense200", "chain"])
def test_edge_shapes(name):
    """Degenerate patterns: one column, no off-diagonal (every front a root
    with na = 0), one dense front (the large-front path end to end), a chain
    of singleton fronts; cholesky / solve / trmm / projected inverse /
    completion against the oracle and the closed forms."""
    import torch
    A = _edge_problems()[name]
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    ref = X.blkval.cpu().numpy().copy()
    assert on.cholesky(symb, ref) == 0
    cp.cholesky(X)
    assert _rel(X.blkval.cpu().numpy(), ref) <= TOL
    n = symb.n
    rng = np.random.default_rng(3)
    Bh = rng.standard_normal((n, 5))
    B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()
    cp.trsm(X, B)
    cp.trsm(X, B, trans="T")
    As = cp.symmetrize(A).toarray()
    assert np.abs(As @ B.cpu().numpy() - Bh).max() <= 1e-12 * np.abs(Bh).max() * max(1, np.abs(As).max())
    cp.trmm(X, B, trans="T")
    cp.trmm(X, B)
    assert _rel(B.cpu().numpy(), Bh) <= 1e-12
    Y = X.copy()
    cp.projected_inverse(Y)
    yr = ref.copy()
    on.projected_inverse(symb, yr)
    assert _rel(Y.blkval.cpu().numpy(), yr) <= TOL
    cp.completion(Y)
    m = _lower_mask(symb)
    assert _rel(Y.blkval.cpu().numpy()[m], ref[m]) <= 1e-11


def test_trsm_offset_and_leading_dimension():
    """trsm / trmm on a column block of a larger array (offsetB, ldB, nrhs as
    cbase.c:1006-1041 reads them): only the addressed entries change."""
    import torch
    A = problems.grid2d(20, seed=5)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    cp.cholesky(X)
    n = symb.n
    ld = n + 7
    rng = np.random.default_rng(1)
    store = rng.standard_normal(ld * 6)
    Bt = torch.as_tensor(store.copy(), device="cuda:0")
    off = ld + 3   # column 1, row 3 of an ld x 6 array: n rows x 3 columns
    cp.trsm(X, Bt, nrhs=3, offsetB=off, ldB=ld)
    got = Bt.cpu().numpy()
    blk = np.asfortranarray(np.stack([store[off + j * ld: off + j * ld + n] for j in range(3)], axis=1))
    on.trsm(symb, X.blkval.cpu().numpy(), blk, trans="N")
    for j in range(3):
        assert _rel(got[off + j * ld: off + j * ld + n], blk[:, j]) <= TOL
    mask = np.ones(ld * 6, dtype=bool)
    for j in range(3):
        mask[off + j * ld: off + j * ld + n] = False
    assert np.array_equal(got[mask], store[mask])


@pytest.mark.parametrize("panel_width", [256, 128])
def test_cholesky_projinv_grid20_outer_panels(panel_width):
    """Cholesky and projected inverse on the 20^3 grid (root front of 608
    columns): every outer panel of the large fronts has several inner
    panels, so the left-looking update inside the panel launches runs with
    64 to panel_width - 64 columns; parity with the oracle's supernodal
    sweeps."""
    print("case grid20 outer panels", panel_width, flush=True)
    A = problems.grid3d(20, seed=2)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    dev = X.blkval.device
    symb._plans[str(dev)] = numeric.Plan(symb, dev, panel_width=panel_width)
    ref = X.blkval.cpu().numpy().copy()
    assert on.cholesky(symb, ref) == 0
    cp.cholesky(X)
    e = _rel(X.blkval.cpu().numpy(), ref)
    print("cholesky rel err %.3e" % e, flush=True)
    assert e <= TOL
    on.projected_inverse(symb, ref)
    cp.projected_inverse(X)
    e = _rel(X.blkval.cpu().numpy(), ref)
    print("projected_inverse rel err %.3e" % e, flush=True)
    assert e <= TOL


@pytest.mark.parametrize("nrhs", [1, 7, 16, 17])
def test_trsm_rhs_counts_large_fronts(nrhs):
    """trsm N and T with 1-17 right-hand sides on a grid whose top fronts are
    large (the products with <= 16 columns run as narrow tiles: masked
    columns below 16, edge row tiles, K not a multiple of 16, split K)
    against the oracle's supernodal solve."""
    import torch
    print("case trsm nrhs", nrhs, flush=True)
    A = problems.grid3d(16, seed=4)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    ref = X.blkval.cpu().numpy().copy()
    assert on.cholesky(symb, ref) == 0
    cp.cholesky(X)
    rng = np.random.default_rng(nrhs)
    Bh = rng.standard_normal((symb.n, nrhs))
    for tr in "NT":
        B = torch.as_tensor(np.asfortranarray(Bh)).to("cuda:0").t().contiguous().t()
        Bref = np.asfortranarray(Bh.copy())
        on.trsm(symb, ref, Bref, trans=tr)
        cp.trsm(X, B, trans=tr)
        e = _rel(B.cpu().numpy(), Bref)
        print("trsm %s rel err %.3e" % (tr, e), flush=True)
        assert e <= TOL


def test_trsm_zero_pivot_nonfinite_not_error():
    """A zero pivot in L11 of the root front (a large front whose inverse the
    solve keeps): trsm gives non-finite entries and no error, as the
    reference's dtrsm does; the inverse is formed again on the device when
    L11 changes, first to the singular block and then back."""
    import torch
    A = problems.grid3d(12, seed=6)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    cp.cholesky(X)
    k = symb.Nsn - 1
    assert symb.snptr[k + 1] - symb.snptr[k] > 64
    rng = np.random.default_rng(2)
    B0 = torch.as_tensor(np.asfortranarray(rng.standard_normal((symb.n, 3)))).to("cuda:0")
    B0 = B0.t().contiguous().t()
    outs = []
    for rep in range(3):  # eager, then the captured graph
        B = B0.clone().t().contiguous().t()
        cp.trsm(X, B)
        outs.append(B.cpu().numpy())
    assert np.isfinite(outs[0]).all()
    assert all(np.array_equal(o, outs[0]) for o in outs)
    p = int(symb.blkptr[k])
    d = X.blkval[p].item()
    X.blkval[p] = 0.0
    B = B0.clone().t().contiguous().t()
    cp.trsm(X, B)
    assert not np.isfinite(B.cpu().numpy()).all()
    X.blkval[p] = d
    B = B0.clone().t().contiguous().t()
    cp.trsm(X, B)
    assert np.array_equal(B.cpu().numpy(), outs[0])


@pytest.mark.parametrize("name", ["grid2d_64", "grid3d_12", "ref31"])
def test_factor_export_on_gpu(name):
    """spmatrix() of a GPU factor (L21 = B21 L11 on the GPU) against the same
    export done on the host from a copy of the blocks, and against the dense
    Cholesky factor of P A P^T."""
    print("case factor export", name, flush=True)
    A, symb, X = _setup(name, 64)
    cp.cholesky(X)
    L_gpu = X.spmatrix().toarray()
    Xh = cp.cspmatrix(symb, blkval=X.blkval.cpu(), factor=True)
    L_host = Xh.spmatrix().toarray()
    assert _rel(L_gpu, L_host) <= 1e-14
    p = np.asarray(symb.p)
    Ad = cp.symmetrize(A).toarray()[np.ix_(p, p)]
    e = _rel(L_gpu, np.linalg.cholesky(Ad))
    print("factor export vs dense Cholesky %.3e" % e, flush=True)
    assert e <= 1e-12
