"""Many independent chordal factorizations at once.

The instances are laid out as one block-diagonal matrix, ordered block by
block (each instance with its own fill-reducing permutation), so the
elimination tree is a forest with one tree per connected component of each
instance.  The level schedule of the plan then puts the fronts of all
instances that sit at the same height into one launch, and the sweeps replay
as one HIP graph: a batch of uneven instances costs one launch sequence as
long as the tallest tree, not one per instance.

``BatchSymbolic(As, p=amd.order)`` builds the joint symbolic object;
``assemble(As)`` the joint cspmatrix on the GPU; ``cholesky``, ``llt``,
``projected_inverse``, ``trsm`` etc. apply to it unchanged; ``logdet(L)``
returns one log-determinant per instance; ``split(x)`` cuts a joint vector
(original ordering) into the instances' pieces.
"""
import numpy as np
import scipy.sparse as sp

from . import amd as _amd
from .symbolic import symbolic, cspmatrix
from . import numeric, _native

__all__ = ["BatchSymbolic"]


class BatchSymbolic(object):
    def __init__(self, As, p=_amd.order, merge_function=None):
        As = [sp.csc_matrix(sp.tril(A)) for A in As]
        self.sizes = np.array([A.shape[0] for A in As], dtype=np.int64)
        self.offsets = np.zeros(len(As) + 1, dtype=np.int64)
        np.cumsum(self.sizes, out=self.offsets[1:])
        perms = []
        for A, off in zip(As, self.offsets[:-1]):
            if p is None:
                q = np.arange(A.shape[0], dtype=np.int64)
            elif callable(p):
                q = np.asarray(p(A), dtype=np.int64)
            else:
                raise ValueError("p must be None or an ordering routine")
            perms.append(q + off)
        self.A = sp.block_diag(As, format="csc")
        self.symb = symbolic(self.A, p=np.concatenate(perms) if perms else None,
                             merge_function=merge_function)
        inst = np.repeat(np.arange(len(As), dtype=np.int64), self.sizes)
        self.colinst = inst[np.asarray(self.symb.p)]   # instance of each reordered column
        self.count = len(As)

    def assemble(self, As=None, device="cuda:0"):
        """Joint cspmatrix of the instances' values (default: the matrices
        the object was built from)."""
        A = self.A if As is None else sp.block_diag([sp.tril(a) for a in As], format="csc")
        return cspmatrix(self.symb, device=device) + A

    def logdet(self, L):
        """log det of every instance from the joint factor L (2 sum of the
        log diagonal entries of each instance's columns)."""
        import torch
        if not (isinstance(L, cspmatrix) and L.is_factor is True):
            raise ValueError("L must be a cspmatrix factor")
        plan, b = numeric._plan(L)
        part = torch.empty(self.symb.n, dtype=torch.float64, device=b.device)
        rc = _native.lib().chp_logdiag(plan.handle, b.data_ptr(), part.data_ptr(),
                                       numeric._stream(b))
        numeric._raise_rc(rc, "logdiag")
        idx = torch.as_tensor(self.colinst, device=b.device)
        out = torch.zeros(self.count, dtype=torch.float64, device=b.device)
        out.index_add_(0, idx, part)
        return 2.0 * out.cpu().numpy()

    def split(self, x):
        """Instance pieces of a joint vector / matrix in the original ordering."""
        return [x[self.offsets[i]:self.offsets[i + 1]] for i in range(self.count)]
