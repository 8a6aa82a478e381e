"""Per-plan autotuning by measurement.

The level schedule has five parameters that trade launch count against
per-launch efficiency, and their best values depend on the pattern: how
the large fronts of the cholesky sweep are factored (``chol_dag``: 1 one
persistent tile-DAG launch per level, 2 panel and GEMM launches, the
default), the
largest clique order handled by the LDS-resident small-front kernels
(``small_max``; fronts above it use the tiled large-front kernels) and the
LDS such a workgroup may take (``lds_cap``, up to the device's per-block
limit: 160 KB on MI355X), the outer panel width of the large-front
factorization (``panel_width``: the K of its trailing Schur updates), and
the split-K cap of the GEMM tasks (``splitk_max``).

``tune(X)`` times the requested sweeps on the caller's own matrix for each
candidate (coordinate descent: the large-front cholesky, then small_max
with its LDS cap, then the panel width, then split-K), keeps the fastest plan on the symbolic object, so
every later call on that pattern uses it, and returns the measurements.
Every candidate produces the same factor up to rounding: the parameters
change only how the work is cut into launches.
"""
import time

import numpy as np

from . import numeric
from .symbolic import cspmatrix

__all__ = ["tune"]

SMALL = [(32, 65536), (64, 65536), (96, 163840), (128, 163840)]
PANEL = [128, 192, 256, 512]
SPLITK = [1, 16]
DAG = [1, 2]


def _time_sweeps(symb, plan, X0, sweeps, warmup, reps):
    import torch
    key = str(X0.blkval.device)
    symb._plans[key] = plan
    X = cspmatrix(symb, blkval=torch.empty_like(X0.blkval))

    def once():
        numeric.copy_blkval(X.blkval, X0.blkval)
        X.is_factor = False
        for s in sweeps:
            if s == "cholesky":
                numeric.cholesky(X)
            elif s == "projected_inverse":
                numeric.projected_inverse(X)
            elif s == "llt":
                numeric.llt(X)
            else:
                raise ValueError("unknown sweep %r" % s)
    # the first calls of a sweep choose between eager launches and graph
    # replay (four calls), so they are not timed
    for _ in range(warmup):
        once()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        once()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def tune(X, sweeps=("cholesky", "projected_inverse"), warmup=5, reps=5):
    """Choose the plan parameters of X's pattern by timing ``sweeps`` on X
    (a cspmatrix on the GPU, not a factor; it is not modified).  Installs
    the fastest plan on X.symb for X's device and returns
    {"best": params, "best_ms": ms, "default_ms": ms, "trials": [(params, ms)]}."""
    if not (isinstance(X, cspmatrix) and X.is_factor is False):
        raise ValueError("X must be a cspmatrix")
    numeric._check_blkval(X)
    symb = X.symb
    dev = X.blkval.device
    trials = []
    cache = {}

    def run(p):
        k = tuple(sorted(p.items()))
        if k not in cache:
            plan = numeric.Plan(symb, dev, small_max=p["small_max"], panel_width=p["panel_width"],
                                splitk_max=p["splitk_max"], lds_cap=p["lds_cap"],
                                chol_dag=p["chol_dag"])
            p = plan.params()     # the values in force (LDS capped by the device)
            ms = _time_sweeps(symb, plan, X, sweeps, warmup, reps)
            cache[k] = (ms, plan, p)
            trials.append((p, round(ms, 4)))
        return cache[k]

    best = {"small_max": 64, "lds_cap": 65536, "panel_width": 256, "splitk_max": 16,
            "chol_dag": 2}
    default_ms = run(best)[0]
    axes = ("dag", "small", "panel", "splitk") if "cholesky" in sweeps else ("small", "panel", "splitk")
    for axis in axes:
        if axis == "dag":
            cands = [dict(best, chol_dag=d) for d in DAG]
        elif axis == "small":
            cands = [dict(best, small_max=s, lds_cap=l) for s, l in SMALL]
        elif axis == "panel":
            cands = [dict(best, panel_width=w) for w in PANEL]
        else:
            cands = [dict(best, splitk_max=k) for k in SPLITK]
        best = min(cands, key=lambda p: run(p)[0])
    ms, plan, params = run(best)
    symb._plans[str(dev)] = plan
    # drop the other candidates' device memory
    for k, v in list(cache.items()):
        if v[1] is not plan:
            del cache[k]
    return {"best": params, "best_ms": round(ms, 4), "default_ms": round(default_ms, 4),
            "sweeps": list(sweeps), "trials": trials}
