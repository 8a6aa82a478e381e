"""Pins of the oracle's Cholesky adjoint sweep (oracle/numeric.py
cholesky_adjoint) on the CPU:

* against torch autograd through an independent forward: the multifrontal
  Cholesky of cholesky.c:26-82 written with torch operations (dense Cholesky
  and triangular solves per front, extend-add by index_put), whose exact
  reverse-mode gradient torch derives by itself;
* for a log-determinant loss (gradient 2 / L_jj on the factor's diagonal)
  the adjoint is the projected inverse P(X^{-1}) (projected_inverse.c),
  a closed form independent of the sweep;
* central differences of a random linear functional of the factor.
"""
import numpy as np
import pytest
import torch

import chompack_amd as cp
from chompack_amd import problems
from oracle import numeric as on

CASES = {
    "ref31": lambda: problems.reference_test_matrix(),
    "grid2d_9": lambda: problems.grid2d(9, seed=0),
    "grid3d_4": lambda: problems.grid3d(4, seed=1),
    "two_cliques": lambda: problems.two_cliques(20, 8, 12, seed=6),
}


def _symb(name):
    A = CASES[name]()
    symb = cp.symbolic(A, p=cp.amd.order if name != "two_cliques" else None)
    X = cp.cspmatrix(symb, device="cpu") + A
    return symb, X.blkval.numpy().copy()


def _torch_cholesky(symb, x):
    """Stored-convention factor of the blkval tensor x, by torch operations
    (differentiable); follows cholesky.c:26-82 front by front."""
    S = on._S(symb)
    out = [None] * len(S.snpost)
    upd = {}
    for k in S.snpost:
        nn, na, nj = S.dims(k)
        blk = x[S.blkptr[k]:S.blkptr[k + 1]].reshape(nn, nj).T     # nj x nn
        low = torch.cat([torch.tril(blk[:nn]), blk[nn:]], 0)
        F = torch.zeros(nj, nj, dtype=x.dtype)
        F = torch.cat([low, torch.zeros(nj, na, dtype=x.dtype)], 1)
        F = F + F.T - torch.diag(torch.diagonal(F))                 # symmetric front
        for c in reversed(S.children(k)):
            r = torch.as_tensor(S.rel(c))
            Uc = upd.pop(c)
            F = F.index_put((r[:, None], r[None, :]), Uc, accumulate=True)
        L11 = torch.linalg.cholesky(F[:nn, :nn])
        if na > 0:
            L21 = torch.linalg.solve_triangular(L11, F[nn:, :nn].T, upper=False).T
            upd[k] = F[nn:, nn:] - L21 @ L21.T
            B21 = torch.linalg.solve_triangular(L11.T, L21.T, upper=True).T
            out[k] = torch.cat([L11, B21], 0)
        else:
            out[k] = L11
    return torch.cat([o.T.reshape(-1) for o in out])


def _lower_mask(symb):
    S = on._S(symb)
    m = np.zeros(int(S.blkptr[-1]), dtype=bool)
    for k in range(len(S.snptr) - 1):
        nn, na, nj = S.dims(k)
        mk = on._blk(m, S, k)
        ii, jj = np.tril_indices(nj, 0, nn)
        mk[ii, jj] = True
    return m


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_torch_autograd(name):
    symb, b0 = _symb(name)
    m = _lower_mask(symb)
    rng = np.random.default_rng(3)
    C = rng.standard_normal(b0.size) * m          # loss = <C, stored factor>
    x = torch.tensor(b0, requires_grad=True)
    Lt = _torch_cholesky(symb, x)
    (Lt * torch.as_tensor(C)).sum().backward()
    want = x.grad.numpy()
    L = b0.copy()
    assert on.cholesky(symb, L) == 0
    assert np.abs(Lt.detach().numpy()[m] - L[m]).max() <= 1e-12 * np.abs(L).max()
    got = on.cholesky_adjoint(symb, L, C) * on.blk_weights(symb)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err <= 1e-11, err


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_of_logdet_is_projected_inverse(name):
    symb, b0 = _symb(name)
    L = b0.copy()
    assert on.cholesky(symb, L) == 0
    S = on._S(symb)
    G = np.zeros_like(L)
    for k in range(len(S.snptr) - 1):
        nn, na, nj = S.dims(k)
        Bk, Gk = on._blk(L, S, k), on._blk(G, S, k)
        Gk[np.arange(nn), np.arange(nn)] = 2.0 / np.diag(Bk[:nn])
    Fb = on.cholesky_adjoint(symb, L, G)
    Y = L.copy()
    assert on.projected_inverse(symb, Y) == 0
    m = _lower_mask(symb)
    err = np.abs(Fb[m] - Y[m]).max() / np.abs(Y[m]).max()
    assert err <= 1e-12, err


def test_adjoint_central_differences():
    symb, b0 = _symb("grid2d_9")
    m = _lower_mask(symb)
    rng = np.random.default_rng(4)
    C = rng.standard_normal(b0.size) * m
    D = rng.standard_normal(b0.size) * m          # a direction in X
    L = b0.copy()
    on.cholesky(symb, L)
    g = on.cholesky_adjoint(symb, L, C) * on.blk_weights(symb)
    h = 1e-6

    def f(t):
        Z = b0 + t * D
        on.cholesky(symb, Z)
        return float((C * Z).sum())
    fd = (f(h) - f(-h)) / (2 * h)
    dd = float((g * D).sum())
    assert abs(fd - dd) <= 1e-7 * abs(dd)
