"""The allocation-free residual entry point (chp_spmv_residual_noalloc)
against chp_spmv_residual with a caller-supplied workspace: bit-identical
results on two calls in a row, the second on a smaller input, and the size
limit rejected on the host before any launch."""
import ctypes

import numpy as np
import pytest

from chompack_amd import _native, mixed, problems

pytestmark = pytest.mark.gpu


def _case(A, nrhs, seed):
    import torch
    S = mixed.SpMat(A, device="cuda:0")
    n = S.n
    rng = np.random.default_rng(seed)
    cm = lambda: torch.empty_strided((n, nrhs), (1, n), dtype=torch.float64, device="cuda:0")
    X, B = cm(), cm()
    X.copy_(torch.as_tensor(rng.standard_normal((n, nrhs))))
    B.copy_(torch.as_tensor(rng.standard_normal((n, nrhs))))
    return S, X, B, cm


def _call(fn, S, X, B, R, nrhs, *work):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    return fn(S.n, S.rowptr.data_ptr(), S.colind.data_ptr(), S.vals.data_ptr(), X.data_ptr(),
              S.n, B.data_ptr(), S.n, R.data_ptr(), S.n, nrhs, *work, stream)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_residual_matches_workspace_residual(order):
    import torch
    lib = _native.lib()
    cases = [(problems.grid3d(20, seed=1), 40, 1), (problems.grid2d(30, seed=0), 7, 2)]
    for A, nrhs, seed in cases:
        print("case noalloc residual n=%d nrhs=%d" % (A.shape[0], nrhs), flush=True)
        S, X, B, cm = _case(A, nrhs, seed)
        R_ref, R_new = cm(), cm()
        work = torch.empty(S.n * nrhs, dtype=torch.float64, device="cuda:0")
        assert _call(lib.chp_spmv_residual, S, X, B, R_ref, nrhs, work.data_ptr()) == 0
        assert _call(lib.chp_spmv_residual_noalloc, S, X, B, R_new, nrhs) == 0
        torch.cuda.synchronize()
        assert torch.equal(R_new, R_ref)
        # and the product itself: R = B - A X with A the symmetric matrix
        As = A + A.T - np.diag(A.diagonal())
        want = B.cpu().numpy() - As @ X.cpu().numpy()
        err = np.abs(R_new.cpu().numpy() - want).max() / np.abs(want).max()
        assert err <= 1e-13
    # a single right-hand side needs no workspace: same result as well
    S, X, B, cm = _case(problems.grid2d(30, seed=0), 1, 3)
    R_ref, R_new = cm(), cm()
    assert _call(lib.chp_spmv_residual, S, X, B, R_ref, 1, None) == 0
    assert _call(lib.chp_spmv_residual_noalloc, S, X, B, R_new, 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(R_new, R_ref)


def test_noalloc_residual_rejects_oversize():
    import torch
    lib = _native.lib()
    S, X, B, cm = _case(problems.grid2d(8, seed=0), 2, 4)
    R = cm()
    stream = torch.cuda.current_stream().cuda_stream
    n_big = _native.RESID_NOALLOC_MAX // 64 + 1       # n * 64 exceeds the workspace
    rc = lib.chp_spmv_residual_noalloc(n_big, S.rowptr.data_ptr(), S.colind.data_ptr(),
                                       S.vals.data_ptr(), X.data_ptr(), n_big, B.data_ptr(),
                                       n_big, R.data_ptr(), n_big, 64, stream)
    assert rc == -1
    assert b"exceeds the fixed workspace" in ctypes.string_at(lib.chp_last_error())
