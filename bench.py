"""Benchmark: supernodal Cholesky + projected inverse on the 40^3 7-point grid.

Workload: SPD matrix on a 40 x 40 x 40 7-point grid (n = 64000), values
a_ij = -U(0,1), diagonal = row sum of |a_ij| + 1 (seed 1), AMD ordering,
default (maximal) supernodes, FP64.  One step = restore the assembled blkval
from a pristine device copy, cholesky(X), projected_inverse(X) -- all inputs
resident in HBM, symbolic analysis excluded.

Prints one JSON line (time-like metric: lower is better).  ``roofline``
describes the dominant kernel, measured with HIP events around every launch
in a separate instrumented step of the same workload; ``cpu_baseline`` times
the reference's own C kernels (oracle/_ref, built from the reference sources
against OpenBLAS) on this host's cores on the full workload.
"""
import argparse
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK = {"fp64_tflops": 78.6, "fp32_tflops": 157.3, "hbm_gbs": 8000.0}
KINDS = ["chol_small", "extadd", "potrf_diag", "panel_factor", "gemm", "extract",
         "projinv_small", "trtri_diag", "llt_small", "solve_fwd", "solve_bwd",
         "copy_top", "splitk_reduce", "diag", "matop", "update_factor", "sgather_fwd",
         "sgather_bwd", "sscatter", "trmm_fwd", "trmm_bwd", "trmm_finish", "completion_chain",
         "projinv_chain", "chol_dag", "fill_int", "gemm_small", "chol_chain", "llt_chain", "uf_setup",
         "uf_step", "gemm_narrow"]
NK = len(KINDS)


def algorithmic_flops(symb):
    """Algorithmic flops per front: cholesky nn^3/3 + 2 na nn^2 + na^2 nn;
    projected inverse nn^3/3 + nn^3 + 2 na^2 nn + 2 nn^2 na."""
    import numpy as np
    nn = np.diff(symb.snptr).astype(float)
    na = np.diff(symb.relptr).astype(float)
    chol = float(np.sum(nn ** 3 / 3 + 2 * na * nn ** 2 + na ** 2 * nn))
    pinv = float(np.sum(nn ** 3 / 3 + nn ** 3 + 2 * na ** 2 * nn + 2 * nn ** 2 * na))
    return chol, pinv


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel), or
    None when the summary is absent."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_grid40_cholesky_projinv.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch %s:" % kernel):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def memory_bytes(symb, small_max=64):
    """Algorithmic HBM bytes of the memory-bound front kernels of one step:
    extend-add into large fronts (12 N (N + 1) B per child of N rows: read the
    child's update triangle, read and write the target entries) and the
    extraction launches: update matrices of the children (read the parent
    triangle, write both triangles: 8 N (N + 1) / 2 + 8 N^2 B) and, in the
    same launches, the large fronts' F21 = -T copies of the projected
    inverse (read T, write F21: 16 na nn B)."""
    import numpy as np
    nn = np.diff(symb.snptr)
    na = np.diff(symb.relptr)
    nj = nn + na
    large = (nj > small_max) | (nj.astype(float) ** 2 * 8 > 64 * 1024)
    ext = 0.0
    chptr, chidx = np.asarray(symb.chptr), np.asarray(symb.chidx)
    for k in np.nonzero(large)[0]:
        for c in chidx[chptr[k]:chptr[k + 1]]:
            N = float(na[c])
            ext += 12.0 * N * (N + 1)
    # k_extract runs for every front with a separator (small and large parents)
    N = na.astype(float)
    xtr = float(np.sum(4.0 * N * (N + 1) + 8.0 * N * N))
    xtr += float(np.sum(16.0 * na[large].astype(float) * nn[large].astype(float)))
    return ext, xtr


def _timed(f, reps):
    """Median wall time (ms) of reps calls of f after one warm-up call."""
    import torch
    f()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(ts)


def batched_section(dev):
    """512 independent 2D 5-point grids, side ~ U{16..96} (seed 2), as one
    block-diagonal factorization: cholesky + per-instance logdet in FP64 and
    with FP32 Schur updates; the logdet error of the mixed path."""
    import numpy as np
    import torch
    import chompack_amd as cp
    from chompack_amd import problems
    As = problems.batch_grids(512, 16, 96, seed=2)
    bs = cp.batch.BatchSymbolic(As)
    X0 = bs.assemble()
    X = cp.cspmatrix(bs.symb, blkval=torch.empty_like(X0.blkval))
    res = {}

    def run(chol):
        def f():
            cp.numeric.copy_blkval(X.blkval, X0.blkval)
            X.is_factor = False
            chol(X)
            res["ld"] = bs.logdet(X)
        return f
    t_default = _timed(run(cp.cholesky), 10)
    tuned = cp.autotune.tune(X0, sweeps=("cholesky",))
    t64 = _timed(run(cp.cholesky), 10)
    ld64 = res["ld"]
    t32 = _timed(run(cp.mixed.cholesky), 10)
    err = float(np.max(np.abs(res["ld"] - ld64) / np.abs(ld64)))
    tbf = _timed(run(lambda Z: cp.mixed.cholesky(Z, schur="bf16x3")), 10)
    errbf = float(np.max(np.abs(res["ld"] - ld64) / np.abs(ld64)))
    return {"instances": 512, "n_total": int(bs.symb.n), "fp64_ms": round(t64, 3),
            "fp64_default_plan_ms": round(t_default, 3), "autotune": tuned,
            "mixed_ms": round(t32, 3), "mixed_vs_fp64_speedup": round(t64 / t32, 3),
            "mixed_logdet_max_rel_err": err,
            "bf16x3_ms": round(tbf, 3), "bf16x3_vs_fp64_speedup": round(t64 / tbf, 3),
            "bf16x3_logdet_max_rel_err": errbf,
            "what": "cholesky + per-instance logdet of the whole batch (mixed: FP32 Schur "
                    "updates; bf16x3: BF16x3 split-precision Schur updates)"}


def rgg_section(dev):
    """Random geometric graph, n = 200000 points in the unit cube, mean degree
    12 (seed 5), AMD: cholesky and a 64-RHS solve in FP64 against FP32 Schur
    updates with FP64 iterative refinement to a 1e-10 relative residual."""
    import numpy as np
    import torch
    import chompack_amd as cp
    from chompack_amd import problems
    A = problems.rgg(200000, 12.0, seed=5)
    symb = cp.symbolic(A, p=cp.amd.order)
    X0 = cp.cspmatrix(symb, device=dev) + A
    X = cp.cspmatrix(symb, blkval=torch.empty_like(X0.blkval))
    rng = np.random.default_rng(0)
    Bh = rng.standard_normal((symb.n, 64))
    B0 = torch.as_tensor(np.asfortranarray(Bh)).to(dev).t().contiguous().t()
    Bs = B0.clone()
    Am = cp.mixed.SpMat(A, device=dev)
    hist = []

    def fp64():
        cp.numeric.copy_blkval(X.blkval, X0.blkval)
        X.is_factor = False
        cp.cholesky(X)
        Bs.copy_(B0)
        cp.trsm(X, Bs)
        cp.trsm(X, Bs, trans="T")

    def mixed(schur):
        def f():
            cp.numeric.copy_blkval(X.blkval, X0.blkval)
            X.is_factor = False
            cp.mixed.cholesky(X, schur=schur)
            _, h = cp.mixed.solve(Am, X, B0, tol=1e-10)
            hist[:] = h
        return f

    def factor_only(schur):
        def f():
            cp.numeric.copy_blkval(X.blkval, X0.blkval)
            X.is_factor = False
            if schur is None:
                cp.cholesky(X)
            else:
                cp.mixed.cholesky(X, schur=schur)
        return f
    t64 = _timed(fp64, 3)
    xs = Bs.cpu().numpy()
    As = cp.symmetrize(A).tocsr()
    r64 = float(np.max(np.linalg.norm(As @ xs - Bh, axis=0) / np.linalg.norm(Bh, axis=0)))
    t32 = _timed(mixed("fp32"), 3)
    h32 = list(hist)
    tbf = _timed(mixed("bf16x3"), 3)
    hbf = list(hist)
    f64, f32, fbf = (_timed(factor_only(m), 3) for m in (None, "fp32", "bf16x3"))
    return {"n": int(symb.n), "nnz_L": int(symb.nnz), "nrhs": 64,
            "fp64_ms": round(t64, 3), "fp64_residual": r64,
            "mixed_ms": round(t32, 3), "mixed_vs_fp64_speedup": round(t64 / t32, 3),
            "mixed_residual_history": h32,
            "bf16x3_ms": round(tbf, 3), "bf16x3_vs_fp64_speedup": round(t64 / tbf, 3),
            "bf16x3_residual_history": hbf,
            "factorization_only_ms": {"fp64": round(f64, 3), "fp32_schur": round(f32, 3),
                                      "bf16x3_schur": round(fbf, 3)},
            "what": "cholesky + solve of 64 N(0,1) right-hand sides, residual <= 1e-10 (mixed: "
                    "FP32 Schur updates + FP64 refinement; bf16x3: BF16x3 split-precision Schur "
                    "updates + FP64 refinement)"}


def dense_completions_section(dev, with_cpu):
    """Dense completions on the 64 x 64 5-point grid (n = 4096, seed 0, AMD):
    psdcompletion of A (dense 4096 x 4096 result) and mrcompletion of
    P(U U^T), U 4096 x 4 N(0,1) (seed 1); the oracle's numpy restatement of
    the reference timed beside it (one run each, 1 thread of numpy/LAPACK
    calls per supernode)."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    import chompack_amd as cp
    from chompack_amd import problems
    A = problems.grid2d(64, seed=0)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device=dev) + A
    U = np.random.default_rng(1).standard_normal((symb.n, 4))
    P = sp.coo_matrix(symb.sparsity_pattern(reordered=False, symmetric=False))
    A2 = sp.csc_matrix((np.einsum("ij,ij->i", U[P.row], U[P.col]), (P.row, P.col)),
                       shape=A.shape)
    X2 = cp.cspmatrix(symb, device=dev) + A2
    res = {}

    def psd():
        res["X"] = cp.psdcompletion(X)

    def mr():
        res["Y"] = cp.mrcompletion(X2)
    t_psd = _timed(psd, 5)
    t_mr = _timed(mr, 5)
    Y = res["Y"].cpu().numpy()
    out = {"n": int(symb.n), "nsn": int(symb.Nsn), "clique_number": int(symb.clique_number),
           "psdcompletion_ms": round(t_psd, 3), "mrcompletion_ms": round(t_mr, 3),
           "mrcompletion_rank": int(Y.shape[1])}
    if with_cpu:
        from oracle import completions as oc
        blk = X.blkval.cpu().numpy()
        t0 = time.perf_counter()
        ref = oc.psdcompletion(symb, blk)
        cpu_psd = (time.perf_counter() - t0) * 1e3
        blk2 = X2.blkval.cpu().numpy()
        t0 = time.perf_counter()
        oc.mrcompletion(symb, blk2)
        cpu_mr = (time.perf_counter() - t0) * 1e3
        Xg = res["X"].cpu().numpy()
        Ad2 = oc.dense_projection(symb, blk2)
        mask = oc.dense_projection(symb, np.ones(blk2.size)) != 0
        out.update({
            "cpu_port_psd_ms": round(cpu_psd, 1), "cpu_port_mr_ms": round(cpu_mr, 1),
            "psd_speedup_vs_cpu": round(cpu_psd / t_psd, 2),
            "mr_speedup_vs_cpu": round(cpu_mr / t_mr, 2),
            "psd_rel_err_vs_port": float(np.abs(Xg - ref).max() / np.abs(ref).max()),
            "mr_pattern_residual": float(np.abs((Y @ Y.T - Ad2)[mask]).max()
                                         / np.abs(Ad2).max())})
    return out


def differentiation_section(dev, with_cpu):
    """-logdet on the 32^3 7-point grid (n = 32768, seed 3): gradient through
    the autograd function (cholesky forward, projected-inverse backward) and
    its central-difference error (step 1e-6) along one random direction; the
    Hessian map (adj=None) on 32 random directions; and the maximum-
    determinant completion of the band-plus-arrow pattern (n = 20000,
    bandwidth 8, 50 arrow columns, seed 4).  The reference's C kernels time
    the Hessian on 4 of the directions (scaled to 32) and the completion."""
    import numpy as np
    import torch
    import chompack_amd as cp
    from chompack_amd import problems
    A = problems.grid3d(32, seed=3)
    symb = cp.symbolic(A, p=cp.amd.order)
    X = cp.cspmatrix(symb, device=dev) + A
    wgt = cp.autograd._grad_weights(symb, X.blkval.device)
    w = torch.where(wgt != 0, 1.0, 0.0).double()
    rng = np.random.default_rng(9)
    res = {}

    def grad():
        Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
        f = -cp.autograd.logdet(Xg)
        f.backward()
        res["G"] = Xg.blkval.grad
    t_grad = _timed(grad, 3)
    U = torch.as_tensor(rng.standard_normal(X.blkval.numel()), device=dev) * w
    dd = float((res["G"] * U).sum())
    h = 1e-6

    def logdiag_at(t):
        Lt = cp.cspmatrix(symb, blkval=X.blkval + t * U)
        cp.cholesky(Lt)
        return cp.numeric.logdiag(Lt)
    fd = -2.0 * float((logdiag_at(h) - logdiag_at(-h)).sum()) / (2 * h)
    # loss = <C, stored factor> through autograd.cholesky: forward + the
    # adjoint sweep
    Cw = torch.as_tensor(rng.standard_normal(X.blkval.numel()), device=dev) * w

    def chol_grad():
        Xg = cp.cspmatrix(symb, blkval=X.blkval.clone().requires_grad_(True))
        (cp.autograd.cholesky(Xg) * Cw).sum().backward()
    t_cgrad = _timed(chol_grad, 3)
    L = X.copy()
    cp.cholesky(L)
    Y = L.copy()
    cp.projected_inverse(Y)
    U0 = [torch.as_tensor(rng.standard_normal(X.blkval.numel()), device=dev) * w for _ in range(32)]
    Us = [cp.cspmatrix(symb, blkval=u.clone()) for u in U0]

    def hess():
        for u, u0 in zip(Us, U0):
            u.blkval.copy_(u0)
        cp.hessian(L, Y, Us, adj=None)
    t_hess = _timed(hess, 3)
    B = problems.band_arrow(20000, 8, 50, seed=4)
    bsymb = cp.symbolic(B, p=None)
    Xb0 = cp.cspmatrix(bsymb, device=dev) + B
    Xb = Xb0.copy()

    def compl():
        cp.numeric.copy_blkval(Xb.blkval, Xb0.blkval)
        Xb.is_factor = False
        cp.completion(Xb)
    t_comp = _timed(compl, 1)
    # gradient of the max log det over completions (llt backward) at the
    # projected inverse of the band-plus-arrow matrix
    Yb = Xb0.copy()
    cp.cholesky(Yb)
    cp.projected_inverse(Yb)

    Xc = Xb0.copy()

    def chol_pinv():
        cp.numeric.copy_blkval(Xc.blkval, Xb0.blkval)
        Xc.is_factor = False
        cp.cholesky(Xc)
        cp.projected_inverse(Xc)
    t_cp = _timed(chol_pinv, 3)

    def maxdet_grad():
        Yg = cp.cspmatrix(bsymb, blkval=Yb.blkval.clone().requires_grad_(True))
        cp.autograd.maxdet(Yg).backward()
    t_md = _timed(maxdet_grad, 1)
    out = {"grid32_n": int(symb.n), "gradient_ms": round(t_grad, 3),
           "cholesky_forward_backward_ms": round(t_cgrad, 3),
           "band_arrow_maxdet_gradient_ms": round(t_md, 1),
           "band_arrow_cholesky_projinv_ms": round(t_cp, 2),
           "gradient_fd_rel_err": abs(dd - fd) / abs(dd), "hessian_32_directions_ms": round(t_hess, 2),
           "band_arrow_n": int(bsymb.n), "band_arrow_supernodes": int(bsymb.Nsn),
           "band_arrow_completion_ms": round(t_comp, 1),
           "what": "-logdet gradient (cholesky + projected-inverse backward) with its central "
                   "difference error; hessian(adj=None) on 32 directions; max-det completion "
                   "(factored updates) of the band-plus-arrow chain"}
    if with_cpu:
        from oracle import ref
        if ref.available():
            threads = min(16, os.cpu_count() or 1)
            ref.set_threads(threads)
            Lh, Yh = L.blkval.cpu().numpy().copy(), Y.blkval.cpu().numpy().copy()
            Uh = [u.cpu().numpy().copy() for u in U0[:4]]
            t0 = time.perf_counter()
            ref.hessian(symb, Lh, Yh, Uh, adj=None)
            t_hc = (time.perf_counter() - t0) * 1e3 * 8
            Hg = Us[0].blkval.cpu().numpy()
            bh = Xb0.blkval.cpu().numpy().copy()
            t0 = time.perf_counter()
            ref.completion(bsymb, bh)
            t_cc = (time.perf_counter() - t0) * 1e3
            gc = Xb.blkval.cpu().numpy()
            # cholesky + projected inverse of the band-plus-arrow chain
            ch = Xb0.blkval.cpu().numpy().copy()
            t0 = time.perf_counter()
            ref.cholesky(bsymb, ch)
            ref.projected_inverse(bsymb, ch)
            t_cpc = (time.perf_counter() - t0) * 1e3
            gpc = Xc.blkval.cpu().numpy()
            out.update({"cpu_band_arrow_cholesky_projinv_ms": round(t_cpc, 1),
                        "band_arrow_cholesky_projinv_speedup_vs_cpu": round(t_cpc / t_cp, 3),
                        "band_arrow_cholesky_projinv_rel_err_vs_reference":
                            float(np.abs(gpc - ch).max() / np.abs(ch).max())})
            out.update({"cpu_reference_threads": threads,
                        "cpu_hessian_32_directions_ms": round(t_hc, 1),
                        "hessian_speedup_vs_cpu": round(t_hc / t_hess, 2),
                        "hessian_rel_err_vs_reference": float(np.abs(Hg - Uh[0]).max() / np.abs(Uh[0]).max()),
                        "cpu_band_arrow_completion_ms": round(t_cc, 1),
                        "band_arrow_completion_speedup_vs_cpu": round(t_cc / t_comp, 3),
                        "band_arrow_completion_rel_err_vs_reference":
                            float(np.abs(gc - bh).max() / np.abs(bh).max()),
                        "cpu_sample": "reference C kernels, %d OpenBLAS threads; Hessian timed on 4 "
                                      "of the 32 directions and scaled by 8" % threads})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)  # covers the sweeps' launch-mode tuning (4 calls)
    ap.add_argument("--side", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default plan parameters instead of autotuning them")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-solve", action="store_true", help="skip the 16-RHS solve section")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the batched (512 grids), random-geometric-graph and dense completion sections")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import chompack_amd as cp
    from chompack_amd import problems, _native

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    A = problems.grid3d(args.side, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    t_symb = time.perf_counter() - t0
    X0 = cp.cspmatrix(symb, device=dev) + A
    pristine = X0.blkval.clone()
    X = cp.cspmatrix(symb, blkval=torch.empty_like(pristine))
    print("# n=%d Nsn=%d nnz(L)=%d clique=%d symbolic %.2fs" %
          (symb.n, symb.Nsn, symb.nnz, symb.clique_number, t_symb), file=sys.stderr, flush=True)
    def step():
        cp.numeric.copy_blkval(X.blkval, pristine)
        X.is_factor = False
        cp.cholesky(X)
        cp.projected_inverse(X)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    print("# timed %d steps: %.3f ms/step" % (args.steps, ms), file=sys.stderr, flush=True)
    gpu_result = X.blkval.cpu().numpy()

    # solve with 16 N(0,1) right-hand sides (L^{-1} then L^{-T}) on the factor
    solve = None
    rng = np.random.default_rng(0)
    if not args.no_solve:
        Bh = rng.standard_normal((symb.n, 16))
        B0 = torch.as_tensor(np.asfortranarray(Bh)).to(dev).t().contiguous().t()
        Lf = cp.cspmatrix(symb, blkval=pristine.clone())
        cp.cholesky(Lf)
        Bs = B0.clone()
        for _ in range(2):
            Bs.copy_(B0)
            cp.trsm(Lf, Bs)
            cp.trsm(Lf, Bs, trans="T")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            Bs.copy_(B0)
            cp.trsm(Lf, Bs)
            cp.trsm(Lf, Bs, trans="T")
        torch.cuda.synchronize()
        solve_ms = (time.perf_counter() - t0) * 1e3 / reps
        xs = Bs.cpu().numpy()
        As = cp.symmetrize(A).tocsr()
        resid = float(np.max(np.linalg.norm(As @ xs - Bh, axis=0) / np.linalg.norm(Bh, axis=0)))
        solve = {"nrhs": 16, "ms": round(solve_ms, 3), "what": "trsm N then T with the factor",
                 "max_rel_residual": resid}

    # products with the factor (trmm N then T, 16 right-hand sides) and the
    # product-form Cholesky factor of I + V diag(a) V^T (n = 40^3, m = 16):
    # factor, then a solve with 16 right-hand sides (N then T)
    trmm = pfsec = None
    if not args.no_solve:
        Ms = B0.clone()
        for _ in range(2):
            Ms.copy_(B0)
            cp.trmm(Lf, Ms)
            cp.trmm(Lf, Ms, trans="T")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            Ms.copy_(B0)
            cp.trmm(Lf, Ms)
            cp.trmm(Lf, Ms, trans="T")
        torch.cuda.synchronize()
        trmm = {"nrhs": 16, "ms": round((time.perf_counter() - t0) * 1e3 / 5, 3),
                "what": "trmm N then T with the factor (B := L^T L B)"}
        trmm_gpu = Ms.cpu().numpy()
        from chompack_amd.pfcholesky import pfchol, pftrsm
        m_pf = 16
        Vh = np.asfortranarray(rng.standard_normal((symb.n, m_pf)) / np.sqrt(symb.n))
        a_pf = rng.uniform(0.5, 2.0, m_pf)
        V0g = torch.as_tensor(Vh).to(dev).t().contiguous().t()
        Vg, Lg, Bg = V0g.clone(), torch.zeros_like(V0g), torch.zeros_like(V0g)
        Ug = B0.clone()

        def pf_step():
            Vg.copy_(V0g)
            pfchol(a_pf, Vg, Lg, Bg)
            Ug.copy_(B0)
            pftrsm(Vg, Lg, Bg, Ug, trans="N")
            pftrsm(Vg, Lg, Bg, Ug, trans="T")

        pf_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            pf_step()
        torch.cuda.synchronize()
        pfsec = {"n": symb.n, "m": m_pf, "nrhs": 16,
                 "ms": round((time.perf_counter() - t0) * 1e3 / 5, 3),
                 "what": "pfchol of I + V diag(a) V^T, then pftrsm N and T on 16 right-hand sides"}
        pf_gpu = Ug.cpu().numpy()

    # mixed precision: FP32 Schur updates + FP64 iterative refinement of the
    # 16-RHS solve to a relative residual <= 1e-10, against the FP64 path
    # (FP64 factorization + the same solve)
    mixed = None
    if not args.no_solve:
        Am = cp.mixed.SpMat(A, device=dev)
        Lm = cp.cspmatrix(symb, blkval=torch.empty_like(pristine))

        def fp64_path():
            cp.numeric.copy_blkval(Lm.blkval, pristine)
            Lm.is_factor = False
            cp.cholesky(Lm)
            Bs.copy_(B0)
            cp.trsm(Lm, Bs)
            cp.trsm(Lm, Bs, trans="T")

        hists = {}

        def mixed_path(schur):
            def f():
                cp.numeric.copy_blkval(Lm.blkval, pristine)
                Lm.is_factor = False
                cp.mixed.cholesky(Lm, schur=schur)
                x, h = cp.mixed.solve(Am, Lm, B0, tol=1e-10)
                hists[schur] = h
            return f

        def factor_path(schur):
            def f():
                cp.numeric.copy_blkval(Lm.blkval, pristine)
                Lm.is_factor = False
                if schur is None:
                    cp.cholesky(Lm)
                else:
                    cp.mixed.cholesky(Lm, schur=schur)
            return f

        res_ms = {}
        for name, fn in (("fp64", fp64_path), ("mixed", mixed_path("fp32")),
                         ("bf16x3", mixed_path("bf16x3")), ("f64", factor_path(None)),
                         ("f32", factor_path("fp32")), ("fbf", factor_path("bf16x3"))):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
            res_ms[name] = (time.perf_counter() - t0) * 1e3 / 5
        mixed = {"what": "cholesky + solve of 16 N(0,1) right-hand sides to max_j ||b_j - A x_j|| / ||b_j|| <= 1e-10",
                 "fp64_ms": round(res_ms["fp64"], 3), "mixed_ms": round(res_ms["mixed"], 3),
                 "mixed_vs_fp64_speedup": round(res_ms["fp64"] / res_ms["mixed"], 3),
                 "mixed_residual_history": hists.get("fp32"),
                 "mixed": "FP32 MFMA Schur updates (k_gemm_tiles32); diagonal blocks, panels, solves and residuals FP64",
                 "bf16x3_ms": round(res_ms["bf16x3"], 3),
                 "bf16x3_vs_fp64_speedup": round(res_ms["fp64"] / res_ms["bf16x3"], 3),
                 "bf16x3_residual_history": hists.get("bf16x3"),
                 "bf16x3": "BF16x3 split-precision Schur updates (k_gemm_tiles_bf3, FP32 accumulation)",
                 "factorization_only_ms": {"fp64": round(res_ms["f64"], 3),
                                           "fp32_schur": round(res_ms["f32"], 3),
                                           "bf16x3_schur": round(res_ms["fbf"], 3)}}

    # instrumented step: HIP events around every launch of both sweeps
    plan = symb._plan(dev)
    L = _native.lib()
    L.chp_profile(plan.handle, 1)
    step()
    torch.cuda.synchronize()
    kms = np.zeros(NK)
    kcnt = np.zeros(NK, dtype=np.int64)
    kfl = np.zeros(NK)
    L.chp_profile_read(plan.handle, _native.ptr(kms), _native.ptr(kcnt), _native.ptr(kfl), NK)
    nrows = int(L.chp_profile_launches(plan.handle, None, 0))
    rows = np.zeros(4 * max(nrows, 1))
    L.chp_profile_launches(plan.handle, _native.ptr(rows), nrows)
    rows = rows.reshape(-1, 4)[:nrows]
    L.chp_profile(plan.handle, 0)
    g = rows[rows[:, 0] == 4]
    gemm_classes = {}
    for lo, hi in [(1, 16), (17, 64), (65, 256), (257, 1024), (1025, 1 << 30)]:
        m = (g[:, 1] >= lo) & (g[:, 1] <= hi)
        if m.any():
            gemm_classes["tiles %d-%d" % (lo, min(hi, int(g[:, 1].max())))] = {
                "launches": int(m.sum()), "ms": round(float(g[m, 3].sum()), 3),
                "tflops": round(float(g[m, 2].sum() / (g[m, 3].sum() * 1e-3) / 1e12), 2)}
    comps = {KINDS[i]: {"ms": round(float(kms[i]), 4), "launches": int(kcnt[i])}
             for i in range(NK) if kcnt[i]}
    dom = int(np.argmax(kms))
    gemm_tf = kfl[4] / (kms[4] * 1e-3) / 1e12 if kms[4] > 0 else 0.0
    roof = {"kernel": "k_gemm_tiles + k_gemm_tiles_lat (FP64 MFMA Schur/projected-inverse tile launches)",
            "bound": "mfma", "achieved": round(gemm_tf, 3), "peak": PEAK["fp64_tflops"],
            "unit": "TFLOP/s", "frac": round(gemm_tf / PEAK["fp64_tflops"], 4),
            "traffic": None, "launches": int(kcnt[4]),
            "avg_launch_us": round(float(kms[4]) * 1e3 / max(int(kcnt[4]), 1), 3),
            "algorithmic_flops_per_step": float(kfl[4]),
            "dominant_kind_by_time": KINDS[dom]}
    gb = np.zeros(4)
    L.chp_plan_gemm_bytes(plan.handle, _native.ptr(gb))
    if args.side == 40:
        roof["traffic"] = profiled_traffic("gemm_kind")
        roof["traffic_source"] = "profiles/hbm_traffic_grid40_cholesky_projinv.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-kernel calibrated)"
    roof["algorithmic_bytes_per_launch"] = float((gb[0] + gb[1]) / max(gb[2] + gb[3], 1))

    # plan parameters chosen by measurement on this matrix, reported beside
    # the default plan's step (the step above and its rocprofv3 summaries use
    # the default plan, so a profile of this command without tuning matches)
    tuned = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuned = cp.autotune.tune(cp.cspmatrix(symb, blkval=pristine))
        tuned["tune_s"] = round(time.perf_counter() - t0, 2)
        print("# autotune: %s %.3f ms (default %.3f ms)" % (tuned["best"], tuned["best_ms"],
              tuned["default_ms"]), file=sys.stderr, flush=True)

    fl_chol, fl_pinv = algorithmic_flops(symb)
    ext_b, xtr_b = memory_bytes(symb)
    hbm = {}
    for kind, nbytes in (("extadd", ext_b), ("extract", xtr_b)):
        i = KINDS.index(kind)
        if kcnt[i] and kms[i] > 0:
            gbs = nbytes / (kms[i] * 1e-3) / 1e9
            hbm[kind] = {"bytes_per_step": nbytes, "ms": round(float(kms[i]), 4),
                         "achieved_GBs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / PEAK["hbm_gbs"], 4)}
    # whole-path roofline: each kernel class at its own bound, summed (the
    # classes are separate launches): all algorithmic flops at the FP64
    # matrix peak, the extend-add and extract bytes at the HBM peak
    t_flops = (fl_chol + fl_pinv) / (PEAK["fp64_tflops"] * 1e12) * 1e3
    t_bytes = (ext_b + xtr_b) / (PEAK["hbm_gbs"] * 1e9) * 1e3
    whole = {"roofline_time_ms": round(t_flops + t_bytes, 4), "measured_ms": round(ms, 4),
             "frac": round((t_flops + t_bytes) / ms, 4),
             "flops_floor_ms": round(t_flops, 4), "bytes_floor_ms": round(t_bytes, 4),
             "formula": "sum over kernel classes of max(flops / 78.6 TFLOP/s, bytes / 8 TB/s): "
                        "cholesky + projected-inverse algorithmic flops, extend-add + extract "
                        "algorithmic bytes"}
    out = {
        "metric": "Cholesky+projinv wall time, 40³ grid; mixed vs FP64 speedup at 1e-10 resid",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D 7-point grid %d^3 (n=%d), seed 1, AMD ordering, cholesky + "
                               "projected_inverse, FP64" % (args.side, symb.n),
                   "n": symb.n, "nnz_L": symb.nnz, "supernodes": symb.Nsn,
                   "clique_number": symb.clique_number},
        "roofline": roof,
        "whole_path_roofline": whole,
        "autotune": tuned,
        "effective_tflops": round((fl_chol + fl_pinv) / (ms * 1e-3) / 1e12, 3),
        "algorithmic_flops": {"cholesky": fl_chol, "projected_inverse": fl_pinv},
        "components_instrumented_ms": comps,
        "hbm_kernels": hbm,
        "solve_16rhs": solve,
        "batched_512_grids": None,
        "rgg_200k": None,
        "trmm_16rhs": trmm,
        "pfcholesky": pfsec,
        "mixed_vs_fp64": mixed,
        "gemm_by_size": gemm_classes,
        "gemm_top": [{"tiles": int(r[1]), "gflop": round(r[2] / 1e9, 3), "ms": round(r[3], 4),
                      "tflops": round(r[2] / (r[3] * 1e-3) / 1e12, 2),
                      "k_equiv": int(r[2] / max(r[1], 1) / (2 * 64 * 64))}
                     for r in sorted(g.tolist(), key=lambda r: -r[3])[:15]],
    }

    if not args.no_configs:
        for key, fn in (("batched_512_grids", batched_section), ("rgg_200k", rgg_section)):
            try:
                out[key] = fn(dev)
            except Exception as e:  # a failed side section must not lose the main line
                out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("# %s done" % key, file=sys.stderr, flush=True)
        try:
            out["differentiation_grid32_band_arrow"] = differentiation_section(dev, not args.no_cpu)
        except Exception as e:
            out["differentiation_grid32_band_arrow"] = {"error": "%s: %s" % (type(e).__name__, e)}
        print("# differentiation_grid32_band_arrow done", file=sys.stderr, flush=True)
        try:
            out["dense_completions_grid64"] = dense_completions_section(dev, not args.no_cpu)
        except Exception as e:
            out["dense_completions_grid64"] = {"error": "%s: %s" % (type(e).__name__, e)}
        print("# dense_completions_grid64 done", file=sys.stderr, flush=True)

    if not args.no_cpu:
        from oracle import ref
        if ref.available():
            threads = min(16, os.cpu_count() or 1)
            ref.set_threads(threads)
            b0 = pristine.cpu().numpy()
            times = []
            for it in range(args.cpu_runs + 1):
                b = b0.copy()
                t0 = time.perf_counter()
                ref.cholesky(symb, b)
                ref.projected_inverse(symb, b)
                dt = time.perf_counter() - t0
                if it > 0:
                    times.append(dt)
            cpu_ms = statistics.median(times) * 1e3
            err = float(np.abs(gpu_result - b).max() / np.abs(b).max())
            # the same workload on one host thread (one run)
            ref.set_threads(1)
            b1 = b0.copy()
            t0 = time.perf_counter()
            ref.cholesky(symb, b1)
            ref.projected_inverse(symb, b1)
            cpu1_ms = (time.perf_counter() - t0) * 1e3
            ref.set_threads(threads)
            out["cpu_baseline"] = {"value": round(cpu_ms, 2), "unit": "ms", "cores": threads,
                                   "kind": "reference",
                                   "sample": "full %d^3 workload, cholesky + projected_inverse "
                                             "with the reference C kernels (OpenBLAS, %d threads: "
                                             "the host share of one GPU; %d CPUs visible), median "
                                             "of %d runs after 1 warm-up; 1 thread: one run" %
                                             (args.side, threads, os.cpu_count() or 0,
                                              args.cpu_runs),
                                   "value_1_thread": round(cpu1_ms, 2),
                                   "cpus_visible": os.cpu_count()}
            out["speedup_vs_cpu"] = round(cpu_ms / ms, 2)
            out["speedup_vs_cpu_1_thread"] = round(cpu1_ms / ms, 2)
            out["parity_rel_err_vs_reference"] = err
            if solve is None:
                out["cpu_baseline"]["sample"] += "; solve section skipped"
            Lr = b0.copy() if solve is not None else None
            if Lr is not None:
                ref.cholesky(symb, Lr)
            st = []
            for it in range(args.cpu_runs + 1 if solve is not None else 0):
                Br = np.asfortranarray(Bh.copy())
                t0 = time.perf_counter()
                ref.trsm(symb, Lr, Br, trans="N")
                ref.trsm(symb, Lr, Br, trans="T")
                if it > 0:
                    st.append(time.perf_counter() - t0)
            if solve is not None:
                solve["cpu_reference_ms"] = round(statistics.median(st) * 1e3, 3)
                solve["speedup_vs_cpu"] = round(solve["cpu_reference_ms"] / solve["ms"], 2)
                solve["rel_err_vs_reference"] = float(np.abs(xs - Br).max() / np.abs(Br).max())
            if trmm is not None:
                # the reference's trmm is Python (pybase/trmm.py): its NumPy
                # restatement is the CPU path, one run
                from oracle import numeric as on
                Mr = np.asfortranarray(Bh.copy())
                t0 = time.perf_counter()
                on.trmm(symb, Lr, Mr, trans="N")
                on.trmm(symb, Lr, Mr, trans="T")
                trmm["cpu_port_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
                trmm["speedup_vs_cpu"] = round(trmm["cpu_port_ms"] / trmm["ms"], 2)
                trmm["rel_err_vs_port"] = float(np.abs(trmm_gpu - Mr).max() / np.abs(Mr).max())
            if pfsec is not None:
                pt = []
                for it in range(args.cpu_runs + 1):
                    Vr = Vh.copy(order="F")
                    Lp = np.zeros_like(Vr, order="F")
                    Bp = np.zeros_like(Vr, order="F")
                    Ur = np.asfortranarray(Bh.copy())
                    t0 = time.perf_counter()
                    assert ref.pf_factor(a_pf, Vr, Lp, Bp) == 0
                    ref.pf_solve(Vr, Lp, Bp, Ur, trans="N")
                    ref.pf_solve(Vr, Lp, Bp, Ur, trans="T")
                    if it > 0:
                        pt.append(time.perf_counter() - t0)
                pfsec["cpu_reference_ms"] = round(statistics.median(pt) * 1e3, 3)
                pfsec["cpu_kind"] = "reference (pfcholesky.c, 1 thread: sequential row loops)"
                pfsec["speedup_vs_cpu"] = round(pfsec["cpu_reference_ms"] / pfsec["ms"], 2)
                pfsec["rel_err_vs_reference"] = float(np.abs(pf_gpu - Ur).max() / np.abs(Ur).max())
        else:
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
