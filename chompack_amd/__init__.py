"""chompack_amd -- MI355X-native supernodal multifrontal chordal matrix library.

Drop-in for the numeric hot path of CHOMPACK (src/python/__init__.py:24-44):
``symbolic``/``cspmatrix`` objects with the reference's block storage, and
``cholesky``, ``llt``, ``projected_inverse``, ``completion``, ``hessian``,
``trsm``, ``trmm``, ``dot``, the product-form factor and the dense
completions running as hand-written HIP kernels on gfx950 through the
C-ABI in include/chompack_amd.h.  Symbolic analysis (AMD, elimination tree,
supernodes, relative indices) runs on the host and is integer bit-exact.
"""
from .symbolic import symbolic, cspmatrix, merge_size_fill, peo
from .misc import tril, triu, symmetrize, perm, eye, lmerge, \
    frontal_add_update, frontal_get_update
from . import amd

from .numeric import cholesky, llt, projected_inverse, completion, hessian, trsm, trmm, dot, \
    logdet
from . import autograd, autotune, batch, mixed
from .pfcholesky import pfcholesky
from .orderings import maxcardsearch, maxchord
from .completions import psdcompletion, mrcompletion, edmcompletion

__version__ = "0.1.0"

__all__ = ["symbolic", "cspmatrix", "merge_size_fill", "peo", "cholesky", "llt",
           "completion", "projected_inverse", "hessian", "trsm", "trmm", "dot", "logdet", "autograd",
           "pfcholesky", "maxcardsearch", "maxchord", "psdcompletion", "mrcompletion", "edmcompletion",
           "tril", "triu",
           "symmetrize", "perm", "eye", "lmerge", "amd"]
