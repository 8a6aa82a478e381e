"""ctypes binding of libchompack_amd.so (the C-ABI declared in include/chompack_amd.h).

The library holds both the host-side symbolic routines (etree, counts,
supernodes, embedding, AMD) and the HIP/gfx950 numeric kernels.  There is no
fallback: if the library is missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libchompack_amd.so")

_lib = None

i64 = ctypes.c_int64
i32 = ctypes.c_int
dbl = ctypes.c_double
vp = ctypes.c_void_p
P64 = ctypes.POINTER(ctypes.c_int64)

# CHP_RESID_NOALLOC_MAX of include/chompack_amd.h: the fixed workspace (in
# doubles) of chp_spmv_residual_noalloc, so n * nrhs must not exceed it
RESID_NOALLOC_MAX = 8388608


def lib():
    """Return the loaded library, loading it on first use (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                "libchompack_amd.so not found at %s; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`" % LIBPATH)
        # torch's HIP runtime (libamdhip64 / libhsa-runtime64 shipped with
        # torch) must be loaded first: the library's NEEDED entries then
        # resolve to that same runtime by soname, so device pointers, streams
        # and the device context are shared with torch.  Loaded the other
        # way round, the process would hold two HIP runtimes.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIBPATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    # host symbolic routines
    L.chp_etree.argtypes = [i64, vp, vp, vp]
    L.chp_etree.restype = None
    L.chp_post_order.argtypes = [i64, vp, vp]
    L.chp_post_order.restype = None
    L.chp_counts.argtypes = [i64, vp, vp, vp, vp, vp]
    L.chp_counts.restype = None
    L.chp_supernodes.argtypes = [i64, vp, vp, vp, vp, vp, vp]
    L.chp_supernodes.restype = i64
    L.chp_embed.argtypes = [i64, vp, vp, vp, vp, vp, vp]
    L.chp_embed.restype = i64
    L.chp_relative_idx.argtypes = [i64, vp, vp, vp, vp, vp, vp]
    L.chp_relative_idx.restype = i64
    L.chp_lmerge.argtypes = [vp, vp, i64, i64, vp]
    L.chp_lmerge.restype = i64
    L.chp_amd_order.argtypes = [i64, vp, vp, vp]
    L.chp_amd_order.restype = i32
    # device plan and numeric sweeps (declared when present)
    _opt(L, "chp_version", [], i32)
    _opt(L, "chp_plan_create", [i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, i32], vp)
    _opt(L, "chp_plan_destroy", [vp], None)
    _opt(L, "chp_plan_create_params", [i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32], vp)
    _opt(L, "chp_plan_params", [vp, vp], i32)
    _opt(L, "chp_update_layout", [i64, vp, vp, vp, vp, vp, vp], i64)
    _opt(L, "chp_plan_info", [vp, vp], i32)
    _opt(L, "chp_cholesky", [vp, vp, vp, vp], i32)
    _opt(L, "chp_llt", [vp, vp, vp], i32)
    _opt(L, "chp_projected_inverse", [vp, vp, vp, vp], i32)
    _opt(L, "chp_trsm", [vp, vp, vp, i32, i32, dbl, i32, vp], i32)
    _opt(L, "chp_trmm", [vp, vp, vp, i32, i32, dbl, i32, vp], i32)
    _opt(L, "chp_maxcardsearch", [i64, vp, vp, i64, vp], None)
    _opt(L, "chp_maxchord", [i64, vp, vp, i64, vp, vp, vp, vp], i64)
    _opt(L, "chp_pf_factor", [i64, i32, vp, vp, i32, vp, i32, vp, i32, vp], i32)
    _opt(L, "chp_pf_solve", [i64, i32, vp, i32, vp, i32, vp, i32, vp, i32, i32, i32, vp], i32)
    _opt(L, "chp_pf_product", [i64, i32, vp, i32, vp, i32, vp, i32, vp, i32, i32, i32, vp], i32)
    _opt(L, "chp_psdcompletion", [i64, i64, vp, vp, vp, vp, vp, vp, i64, dbl, vp], i32)
    _opt(L, "chp_edmcompletion", [i64, i64, vp, vp, vp, vp, vp, vp, i64, dbl, vp], i32)
    _opt(L, "chp_mrcompletion", [i64, i64, vp, vp, vp, vp, vp, vp, vp, i64, i32, vp, vp], i32)
    _opt(L, "chp_dense_permute", [vp, i64, vp, i64, i64, i32, vp, i64, vp], i32)
    _opt(L, "chp_dot", [vp, vp, vp, vp, vp], i32)
    _opt(L, "chp_logdet", [vp, vp, vp, vp], i32)
    _opt(L, "chp_logdiag", [vp, vp, vp, vp], i32)
    _opt(L, "chp_copy", [vp, vp, i64, vp], i32)
    _opt(L, "chp_factor_export", [vp, vp, vp, vp], i32)
    _opt(L, "chp_cholesky_mixed", [vp, vp, vp, vp], i32)
    _opt(L, "chp_cholesky_lowprec", [vp, vp, i32, vp, vp], i32)
    _opt(L, "chp_spmv_residual", [i64, vp, vp, vp, vp, i32, vp, i32, vp, i32, i32, vp, vp], i32)
    _opt(L, "chp_spmv_residual_noalloc", [i64, vp, vp, vp, vp, i32, vp, i32, vp, i32, i32, vp],
         i32)
    _opt(L, "chp_completion", [vp, vp, i32, vp, vp], i32)
    _opt(L, "chp_cholesky_adjoint", [vp, vp, vp, vp, vp, vp], i32)
    _opt(L, "chp_hessian", [vp, vp, vp, vp, i32, i32, i32, i32, vp, vp], i32)
    _opt(L, "chp_frontal_add_update", [vp, i64, vp, vp, vp, dbl, vp], i32)
    _opt(L, "chp_frontal_get_update", [vp, i64, vp, vp, vp, vp], i32)
    _opt(L, "chp_profile", [vp, i32], i32)
    _opt(L, "chp_set_graphs", [vp, i32], i32)
    _opt(L, "chp_profile_read", [vp, vp, vp, vp, i32], i32)
    _opt(L, "chp_dag_trace", [vp, i32], i64)
    _opt(L, "chp_dag_trace_read", [vp, vp, i64], i64)
    _opt(L, "chp_set_dag_wait_limit", [vp, i32], i32)
    _opt(L, "chp_plan_flops", [vp, vp], i32)
    _opt(L, "chp_plan_gemm_bytes", [vp, vp], i32)
    _opt(L, "chp_profile_launches", [vp, vp, i64], i64)
    _opt(L, "chp_last_error", [], ctypes.c_char_p)


def _opt(L, name, argtypes, restype):
    if hasattr(L, name):
        f = getattr(L, name)
        f.argtypes = argtypes
        f.restype = restype


def ptr(a):
    """Host pointer of a C-contiguous numpy array."""
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def as_i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))
