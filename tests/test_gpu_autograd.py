"""GPU tests of the differentiable factorizations (chompack_amd.autograd)
and the Cholesky adjoint sweep (chp_cholesky_adjoint).

* the adjoint sweep against the oracle's (oracle/numeric.py
  cholesky_adjoint, itself pinned against torch autograd, the projected
  inverse and central differences in tests/test_oracle_adjoint.py), 1e-12
  relative, on the parity problems;
* the gradients of autograd.cholesky (both storage modes), completion and
  maxdet against central differences (step 1e-6, FP64) of the forward along
  random directions, relative error <= 1e-6, at the sizes of the
  differentiation workload (32^3 grid, band-plus-arrow n = 20000) and on
  smaller patterns;
* logdet's gradient through the Cholesky adjoint equals the projected
  inverse (projected_inverse.c), 1e-12.
"""
import numpy as np
import pytest

import chompack_amd as cp
from chompack_amd import autograd, numeric, problems
from oracle import numeric as on
from tests.test_gpu_parity import PROBLEMS, ORDER, _lower_mask, _factor_pair, _rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_cholesky_adjoint_vs_oracle(name):
    import torch
    print("case adjoint", name, flush=True)
    symb, L, _ = _factor_pair(name)
    m = _lower_mask(symb)
    rng = np.random.default_rng(21)
    G = rng.standard_normal(L.size) * m
    ref = on.cholesky_adjoint(symb, L, G)
    Lc = cp.cspmatrix(symb, blkval=torch.as_tensor(L, device="cuda:0"), factor=True)
    Gc = cp.cspmatrix(symb, blkval=torch.as_tensor(G, device="cuda:0"))
    Fb = numeric.cholesky_adjoint(Lc, Gc)
    e = _rel(Fb.blkval.cpu().numpy()[m], ref[m])
    print("adjoint rel err %.3e" % e, flush=True)
    assert e <= TOL


def _grid32():
    import torch
    A = problems.grid3d(32, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device="cuda:0") + A
    w = torch.as_tensor(_lower_mask(symb).astype(float), device="cuda:0")
    return symb, X, w


@pytest.mark.parametrize("recompute", [False, True])
def test_cholesky_gradient_grid32(recompute):
    """loss = <C, stored factor> on the 32^3 grid: backward (the adjoint
    sweep) against the central difference along a random symmetric
    direction; with recompute=True the backward factors X again."""
    import torch
    print("case autograd cholesky grid32", recompute, flush=True)
    symb, X, w = _grid32()
    rng = np.random.default_rng(5)
    C = torch.as_tensor(rng.standard_normal(len(w)), device="cuda:0") * w
    D = torch.as_tensor(rng.standard_normal(len(w)), device="cuda:0") * w
    Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
    Lb = autograd.cholesky(Xg, recompute=recompute)
    (Lb * C).sum().backward()
    dd = float((Xg.blkval.grad * D).sum())
    h = 1e-6

    def f(t):
        Z = cp.cspmatrix(symb, blkval=X.blkval + t * D)
        cp.cholesky(Z)
        return Z.blkval
    fd = float(((f(h) - f(-h)) * C).sum()) / (2 * h)
    print("grid32 cholesky directional derivative %.10e fd %.10e" % (dd, fd), flush=True)
    assert abs(dd - fd) <= 1e-6 * abs(dd)


def test_logdet_through_cholesky_adjoint_is_projected_inverse():
    """d/dX of 2 sum log L_jj through autograd.cholesky equals the logdet
    gradient P(X^{-1}) (weighted), on the 32^3 grid."""
    import torch
    symb, X, w = _grid32()
    Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
    Lb = autograd.cholesky(Xg)
    ld = 2.0 * torch.log(Lb[torch.as_tensor(_diag_positions(symb), device="cuda:0")]).sum()
    ld.backward()
    Y = X.copy()
    cp.cholesky(Y)
    assert abs(float(ld.detach()) - cp.logdet(Y)) <= 1e-12 * abs(float(ld.detach()))
    cp.projected_inverse(Y)
    want = Y.blkval * autograd._grad_weights(symb, Y.blkval.device)
    e = float((Xg.blkval.grad - want).abs().max() / want.abs().max())
    print("logdet via adjoint vs projected inverse %.3e" % e, flush=True)
    assert e <= TOL


def _diag_positions(symb):
    pos = []
    for k in range(symb.Nsn):
        nn = int(symb.snptr[k + 1] - symb.snptr[k])
        nj = nn + int(symb.relptr[k + 1] - symb.relptr[k])
        pos += [int(symb.blkptr[k]) + j * nj + j for j in range(nn)]
    return np.asarray(pos, dtype=np.int64)


def _completable(name):
    """X = P(A^{-1}) of a parity problem: a completable partial matrix whose
    max-det completion is A^{-1} (so L L^T = A)."""
    import torch
    symb, L, Y = _factor_pair(name)
    return symb, cp.cspmatrix(symb, blkval=torch.as_tensor(Y, device="cuda:0"))


@pytest.mark.parametrize("name", ["grid2d_64", "two_cliques", "caterpillar"])
def test_completion_gradient(name):
    """loss = <C, completion factor>: backward (adjoint sweep, then the
    inverse Hessian map) against central differences."""
    import torch
    print("case autograd completion", name, flush=True)
    symb, X = _completable(name)
    w = torch.as_tensor(_lower_mask(symb).astype(float), device="cuda:0")
    rng = np.random.default_rng(6)
    C = torch.as_tensor(rng.standard_normal(len(w)), device="cuda:0") * w
    D = torch.as_tensor(rng.standard_normal(len(w)), device="cuda:0") * w
    D *= 1e-2 * float(X.blkval.abs().max()) / float(D.abs().max())
    Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
    Lb = autograd.completion(Xg)
    (Lb * C).sum().backward()
    dd = float((Xg.blkval.grad * D).sum())
    h = 1e-6

    def f(t):
        Z = cp.cspmatrix(symb, blkval=X.blkval + t * D)
        cp.completion(Z)
        return Z.blkval
    fd = float(((f(h) - f(-h)) * C).sum()) / (2 * h)
    print("completion directional derivative %.10e fd %.10e" % (dd, fd), flush=True)
    assert abs(dd - fd) <= 1e-6 * abs(dd)


def test_maxdet_gradient_band_arrow():
    """The max log det over completions of the band-plus-arrow pattern
    (n = 20000, bandwidth 8, 50 arrow columns, seed 4): gradient (L L^T by
    the llt sweep) against the central difference along a random direction,
    and against the oracle's llt of the oracle's completion factor."""
    import torch
    B = problems.band_arrow(20000, 8, 50, seed=4)
    symb = cp.symbolic(B, p=None)
    X = cp.cspmatrix(symb, device="cuda:0") + B
    # a completable point: the projected inverse of the SPD matrix B
    Y = X.copy()
    cp.cholesky(Y)
    cp.projected_inverse(Y)
    w = torch.as_tensor(_lower_mask(symb).astype(float), device="cuda:0")
    rng = np.random.default_rng(7)
    D = torch.as_tensor(rng.standard_normal(len(w)), device="cuda:0") * w
    D *= 1e-2 * float(Y.blkval.abs().max()) / float(D.abs().max())
    Yg = cp.cspmatrix(symb, blkval=Y.blkval.clone().requires_grad_(True))
    v = autograd.maxdet(Yg)
    v.backward()
    dd = float((Yg.blkval.grad * D).sum())
    h = 1e-6

    def ldiag(t):
        Z = cp.cspmatrix(symb, blkval=Y.blkval + t * D)
        cp.completion(Z)
        return numeric.logdiag(Z)
    # per-column differences keep their digits (the totals are ~1e4)
    fd = -2.0 * float((ldiag(h) - ldiag(-h)).sum()) / (2 * h)
    print("band-arrow maxdet directional derivative %.10e fd %.10e" % (dd, fd), flush=True)
    assert abs(dd - fd) <= 1e-6 * abs(dd)
    # the gradient is (L L^T) on the pattern: here L L^T = B (weighted)
    want = X.blkval * autograd._grad_weights(symb, X.blkval.device)
    e = float(((Yg.blkval.grad - want) * w).abs().max() / (want * w).abs().max())
    print("band-arrow maxdet gradient vs B %.3e" % e, flush=True)
    assert e <= 1e-10
