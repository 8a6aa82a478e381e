"""GPU tests of the tunable plan parameters and the autotuner
(chompack_amd.autotune): every parameter choice gives the oracle's
cholesky, projected inverse and llt (1e-12 relative), including small-front
kernels with more than 64 KB of LDS, persistent launches over
largest-first front lists and the tile-DAG factorization of large fronts
(chol_dag=1, also with every front large); tune() installs its fastest plan
and the factor it produces matches the default plan's.  The DAG launch's
bounded waits: with a wait bound of 0 polls the sweep reports the internal
timeout (the launch still drains), and the plan factors correctly again once
the bound is restored."""
import numpy as np
import pytest

import chompack_amd as cp
from chompack_amd import numeric, problems
from oracle import numeric as on
from tests.test_gpu_parity import _rel

pytestmark = pytest.mark.gpu

TOL = 1e-12

PARAMS = [
    dict(small_max=128, lds_cap=163840),
    dict(small_max=96, lds_cap=163840, panel_width=128),
    dict(small_max=32, panel_width=512, splitk_max=1),
    dict(small_max=64, chol_dag=1),
    dict(small_max=1, chol_dag=1),
]


@pytest.mark.parametrize("params", PARAMS, ids=["small128", "small96_panel128", "small32_panel512_nosplit",
                                                "dag", "dag_all_large"])
@pytest.mark.parametrize("name", ["grid2d_64", "grid3d_12"])
def test_plan_parameters_parity(name, params):
    import torch
    print("case plan params", name, params, flush=True)
    A = problems.grid2d(64, seed=0) if name == "grid2d_64" else problems.grid3d(12, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    plan = numeric.Plan(symb, "cuda:0", **params)
    got = plan.params()
    assert got["small_max"] == params["small_max"]
    symb._plans["cuda:0"] = plan
    X = cp.cspmatrix(symb, device="cuda:0") + A
    ref = X.blkval.cpu().numpy().copy()
    A0 = ref.copy()
    assert on.cholesky(symb, ref) == 0
    cp.cholesky(X)
    assert _rel(X.blkval.cpu().numpy(), ref) <= TOL
    L = X.copy()
    L.is_factor = True
    cp.llt(L)
    assert _rel(L.blkval.cpu().numpy(), A0) <= 1e-12
    on.projected_inverse(symb, ref)
    cp.projected_inverse(X)
    e = _rel(X.blkval.cpu().numpy(), ref)
    print("projected inverse rel err %.3e" % e, flush=True)
    assert e <= TOL


def test_tune_batched_grids():
    """tune() on 64 independent grids (side 16..96) picks a plan by
    measurement and installs it; its factor equals the default plan's."""
    import torch
    As = problems.batch_grids(64, 16, 96, seed=2)
    bs = cp.batch.BatchSymbolic(As)
    X0 = bs.assemble()
    Ld = X0.copy()
    cp.cholesky(Ld)
    res = cp.autotune.tune(X0, sweeps=("cholesky",), warmup=5, reps=3)
    print("tune:", res["best"], res["best_ms"], "default", res["default_ms"], flush=True)
    assert len(res["trials"]) >= 5
    assert res["best_ms"] <= res["default_ms"] * 1.05
    plan = X0.symb._plans[str(X0.blkval.device)]
    assert plan.params() == res["best"]
    Lt = X0.copy()
    cp.cholesky(Lt)
    assert _rel(Lt.blkval.cpu().numpy(), Ld.blkval.cpu().numpy()) <= TOL
    assert np.allclose(bs.logdet(Lt), bs.logdet(Ld), rtol=1e-12)


def test_dag_wait_bound():
    """chol_dag=1 with every front large (grid3d_12): a wait bound of 0 polls
    makes some dependency wait run out (tasks of a level start together), the
    sweep returns the internal timeout error after the launch drains; with
    the default bound restored the same plan factors to the oracle's."""
    from chompack_amd import _native
    A = problems.grid3d(12, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    plan = numeric.Plan(symb, "cuda:0", small_max=1, chol_dag=1)
    symb._plans["cuda:0"] = plan
    X = cp.cspmatrix(symb, device="cuda:0") + A
    ref = X.blkval.cpu().numpy().copy()
    assert on.cholesky(symb, ref) == 0
    lib = _native.lib()
    assert lib.chp_set_dag_wait_limit(plan.handle, 0) == 0
    Y = X.copy()
    with pytest.raises(RuntimeError, match="timed out"):
        cp.cholesky(Y)
    assert lib.chp_set_dag_wait_limit(plan.handle, 1 << 18) == 0
    Y = X.copy()
    cp.cholesky(Y)
    e = _rel(Y.blkval.cpu().numpy(), ref)
    print("after the timeout: cholesky rel err %.3e" % e, flush=True)
    assert e <= TOL
