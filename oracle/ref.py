"""ctypes driver of oracle/_ref/libchompack_ref.so -- the reference's own C
numeric kernels (CHOMPACK src/C/*.c) compiled by oracle/Makefile against
SciPy's LP64 OpenBLAS.

ORACLE / TEST INFRASTRUCTURE ONLY.  Imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
product.  The wrappers reproduce the workspace sizing and argument
conventions of the reference's CPython layer (src/C/cbase.c:187-1077) so the
kernels run exactly as chompack runs them; inputs are the arrays of a
``chompack_amd.symbolic`` object and float64 numpy block arrays.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_ref", "libchompack_ref.so")
_lib = None
_blas = None

c_i64 = ctypes.c_int64
P = ctypes.c_void_p


def available():
    return os.path.exists(LIBPATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref not built (run `make -C oracle`)")
        L = ctypes.CDLL(LIBPATH)
        sweep = [c_i64, c_i64, P, P, P, P, P, P, P, P, P, P, P]
        L.cholesky.argtypes = sweep
        L.cholesky.restype = ctypes.c_int
        L.llt.argtypes = sweep
        L.llt.restype = None
        L.projected_inverse.argtypes = sweep
        L.projected_inverse.restype = ctypes.c_int
        L.completion.argtypes = sweep + [ctypes.c_int]
        L.completion.restype = ctypes.c_int
        L.hessian.argtypes = [c_i64, c_i64, P, P, P, P, P, P, P, P, P, P, P, P, P,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.hessian.restype = ctypes.c_int
        L.trsm.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_double, c_i64, c_i64,
                           P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
        L.trsm.restype = None
        L.dot.argtypes = [P, P, P, P, P, P]
        L.dot.restype = ctypes.c_double
        L.dpftrf.argtypes = [P, P, P, P, P, P, P, P, P]
        L.dpftrf.restype = ctypes.c_int
        L.dpfsv.argtypes = [P, P, P, P, P, P, P, P, P, P]
        L.dpfsv.restype = None
        L.dpfmv.argtypes = [P, P, P, P, P, P, P, P, P, P]
        L.dpfmv.restype = None
        _lib = L
    return _lib


def set_threads(n):
    """Set the OpenBLAS thread count used by the reference kernels."""
    global _blas
    if _blas is None:
        import scipy
        d = os.path.realpath(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs"))
        f = [x for x in os.listdir(d) if x.startswith("libscipy_openblas-") and x.endswith(".so")]
        _blas = ctypes.CDLL(os.path.join(d, f[0]))
    _blas.scipy_openblas_set_num_threads(ctypes.c_int(int(n)))


def _p(a):
    return a.ctypes.data


class _Sym(object):
    """int64 copies of the symbolic arrays the reference kernels read."""

    def __init__(self, symb):
        f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64))
        self.n = symb.n
        self.nsn = symb.Nsn
        self.snpost = f(symb.snpost)
        self.snptr = f(symb.snptr)
        self.snode = f(symb.snode)
        self.relptr = f(symb.relptr)
        self.relidx = f(symb.relidx)
        self.chptr = f(symb.chptr)
        self.chidx = f(symb.chidx)
        self.blkptr = f(symb.blkptr)
        self.sncolptr = f(symb.sncolptr)
        self.p = f(symb.p)
        m = symb.memory
        self.stack_depth = max(int(m["stack_depth"]), 1)
        self.stack_mem = max(int(m["stack_mem"]), 1)
        self.frontal_mem = max(int(m["frontal_mem"]), 1)
        self.stack_solve = max(int(m["stack_solve"]), 1)
        self.cln = max(int(symb.clique_number), 1)

    def args(self):
        return (self.n, self.nsn, _p(self.snpost), _p(self.snptr), _p(self.relptr),
                _p(self.relidx), _p(self.chptr), _p(self.chidx), _p(self.blkptr))

    def work(self):
        fws = np.zeros(self.frontal_mem)
        upd = np.zeros(self.stack_mem)
        usz = np.zeros(self.stack_depth, dtype=np.int64)
        return fws, upd, usz


def _sym(symb):
    s = getattr(symb, "_ref_sym", None)
    if s is None:
        s = _Sym(symb)
        symb._ref_sym = s
    return s


def _blk(blkval):
    assert isinstance(blkval, np.ndarray) and blkval.dtype == np.float64 \
        and blkval.flags["C_CONTIGUOUS"]
    return blkval


def cholesky(symb, blkval):
    """In-place supernodal Cholesky (cbase.c:187-284); returns LAPACK info."""
    s = _sym(symb)
    fws, upd, usz = s.work()
    return lib().cholesky(*s.args(), _p(_blk(blkval)), _p(fws), _p(upd), _p(usz))


def llt(symb, blkval):
    """In-place product L L^T (cbase.c:298-390)."""
    s = _sym(symb)
    fws, upd, usz = s.work()
    lib().llt(*s.args(), _p(_blk(blkval)), _p(fws), _p(upd), _p(usz))
    return 0


def projected_inverse(symb, blkval):
    """In-place projected inverse (cbase.c:404-500); returns LAPACK info."""
    s = _sym(symb)
    fws, upd, usz = s.work()
    return lib().projected_inverse(*s.args(), _p(_blk(blkval)), _p(fws), _p(upd), _p(usz))


def completion(symb, blkval, factored_updates=True):
    """In-place max-det completion (cbase.c:521-621); returns info."""
    s = _sym(symb)
    fws, upd, usz = s.work()
    return lib().completion(*s.args(), _p(_blk(blkval)), _p(fws), _p(upd), _p(usz),
                            int(bool(factored_updates)))


def hessian(symb, lblkval, yblkval, ublkvals, adj=False, inv=False, factored_updates=True):
    """In-place Hessian operator on each array of ``ublkvals``
    (cbase.c:675-857); ``adj=None`` applies the operator and its adjoint."""
    s = _sym(symb)
    fws, upd, usz = s.work()
    arr = (ctypes.c_void_p * (len(ublkvals) + 1))()
    for i, u in enumerate(ublkvals):
        arr[i] = _p(_blk(u))
    arr[len(ublkvals)] = None
    inv_i = 1 if inv else 0
    adj_i = inv_i if adj is None else (1 if adj else 0)
    L = lib()
    info = L.hessian(*s.args(), _p(_blk(lblkval)), _p(_blk(yblkval)), ctypes.cast(arr, P),
                     _p(fws), _p(upd), _p(usz), inv_i, adj_i, int(bool(factored_updates)))
    if adj is None:
        info = L.hessian(*s.args(), _p(lblkval), _p(yblkval), ctypes.cast(arr, P),
                         _p(fws), _p(upd), _p(usz), inv_i, 1 ^ adj_i,
                         int(bool(factored_updates)))
    return info


def trsm(symb, lblkval, B, alpha=1.0, trans="N"):
    """B := alpha L^{-1} B or alpha L^{-T} B for a column-major (Fortran
    ordered) float64 array B with n rows (cbase.c:965-1077)."""
    s = _sym(symb)
    assert B.dtype == np.float64 and B.flags["F_CONTIGUOUS"] and B.shape[0] == s.n
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    ld = ctypes.c_int(s.n)
    fws = np.zeros(s.cln * nrhs)
    upd = np.zeros(s.stack_solve * nrhs)
    usz = np.zeros(s.stack_depth, dtype=np.int64)
    lib().trsm(trans.encode(), nrhs, float(alpha), s.n, s.nsn, _p(s.snpost), _p(s.snptr),
               _p(s.snode), _p(s.relptr), _p(s.relidx), _p(s.chptr), _p(s.chidx),
               _p(s.blkptr), _p(s.p), _p(_blk(lblkval)), B.ctypes.data, ctypes.addressof(ld),
               _p(fws), _p(upd), _p(usz))


def dot(symb, x, y):
    """Trace inner product tr(XY) of two chordal matrices (dot.c)."""
    s = _sym(symb)
    nsn = ctypes.c_int64(s.nsn)
    return lib().dot(ctypes.addressof(nsn), _p(s.snptr), _p(s.sncolptr), _p(s.blkptr),
                     _p(_blk(x)), _p(_blk(y)))


def _ints(*v):
    return [ctypes.c_int(int(x)) for x in v]


def pf_factor(a, V, L, B):
    """dpftrf (pfcholesky.c:52-87) on column-major n x m arrays, in place;
    returns 0 or -1."""
    n, m = V.shape
    for M in (V, L, B):
        assert M.dtype == np.float64 and M.flags["F_CONTIGUOUS"]
    a = np.ascontiguousarray(a, dtype=np.float64)
    cn, cm = _ints(n, m)
    return lib().dpftrf(ctypes.addressof(cn), ctypes.addressof(cm), _p(a), V.ctypes.data,
                        ctypes.addressof(cn), L.ctypes.data, ctypes.addressof(cn), B.ctypes.data,
                        ctypes.addressof(cn))


def _pf_apply(fn, V, L, B, U, trans):
    n, m = V.shape
    assert U.dtype == np.float64 and U.flags["F_CONTIGUOUS"] and U.shape[0] == n
    cn, cm = _ints(n, m)
    tr = ctypes.c_char(trans.encode())
    for c in range(1 if U.ndim == 1 else U.shape[1]):
        fn(ctypes.addressof(cn), ctypes.addressof(cm), ctypes.addressof(tr), V.ctypes.data,
           ctypes.addressof(cn), L.ctypes.data, ctypes.addressof(cn), B.ctypes.data,
           ctypes.addressof(cn), U.ctypes.data + 8 * n * c)


def pf_solve(V, L, B, U, trans="N"):
    """dpfsv (pfcholesky.c:89-112) for every column of U, as cbase.c pftrsm."""
    _pf_apply(lib().dpfsv, V, L, B, U, trans)


def pf_product(V, L, B, U, trans="N"):
    """dpfmv (pfcholesky.c:114-136) for every column of U, as cbase.c pftrmm."""
    _pf_apply(lib().dpfmv, V, L, B, U, trans)
