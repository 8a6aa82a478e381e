"""Writes tests/golden/*.npz: input/output vectors of the reference's own C
numeric kernels (CHOMPACK src/C, built by oracle/Makefile into oracle/_ref)
on small problems, driven with this package's symbolic arrays.

Run in the build container (needs /root/reference and oracle/_ref):
    make -C oracle && python tests/golden/make_golden.py
Each fixture holds the symbolic arrays it was made with, so tests can check
both the symbolic layer (bit-exact) and the numeric sweeps (to rounding).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import chompack_amd as cp  # noqa: E402
from chompack_amd import problems  # noqa: E402
from oracle import ref  # noqa: E402


def make(name, A, p=cp.amd.order, nrhs=2):
    symb = cp.symbolic(A, p=p)
    X = cp.cspmatrix(symb, device="cpu") + A
    b0 = X.blkval.numpy().copy()
    out = {"blkval_A": b0}
    L = b0.copy()
    assert ref.cholesky(symb, L) == 0
    out["blkval_L"] = L.copy()
    Y = L.copy()
    assert ref.projected_inverse(symb, Y) == 0
    out["blkval_Y"] = Y
    P = L.copy()
    ref.llt(symb, P)
    out["blkval_LLT"] = P
    rng = np.random.default_rng(11)
    B = np.asfortranarray(rng.standard_normal((symb.n, nrhs)))
    out["B"] = B.copy()
    BN = B.copy(order="F")
    ref.trsm(symb, L, BN, trans="N")
    out["B_trsmN"] = BN
    BT = B.copy(order="F")
    ref.trsm(symb, L, BT, trans="T")
    out["B_trsmT"] = BT
    out["dot_AY"] = np.array(ref.dot(symb, b0, Y))
    for key in ("p", "snptr", "snpar", "relptr", "relidx", "chptr", "chidx", "blkptr",
                "sncolptr", "snrowidx"):
        out["symb_" + key] = np.asarray(getattr(symb, key))
    out["symb_fill"] = np.array(symb.fill)
    out["symb_memory"] = np.array([symb.memory[k] for k in
                                   ("stack_depth", "stack_mem", "frontal_mem", "stack_solve")])
    out["A_rows"], out["A_cols"] = A.nonzero()
    out["A_vals"] = np.asarray(A[out["A_rows"], out["A_cols"]]).ravel()
    out["n"] = np.array(symb.n)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "n", symb.n, "Nsn", symb.Nsn, "nnz", symb.nnz)


if __name__ == "__main__":
    make("ref31_amd", problems.reference_test_matrix())
    make("grid2d_16_amd", problems.grid2d(16, seed=0))
    make("grid3d_6_natural", problems.grid3d(6, seed=1), p=None)
