"""Update-arena layout (host): every pair of update slots whose level
intervals [level(k), level(parent(k))] meet occupies disjoint memory, the
arena is no longer than sum(na^2), and each slot fits inside it.  Checked
by brute force per level (every slot live at the level, sorted by offset,
must not overlap its successor)."""
import numpy as np
import pytest

import chompack_amd as cp
from chompack_amd import _native, problems


def _layout(symb):
    n = symb.Nsn
    a = [_native.as_i64(x) for x in (symb.snpost, symb.relptr, symb.chptr, symb.chidx)]
    upd = np.zeros(n, dtype=np.int64)
    lev = np.zeros(n, dtype=np.int64)
    ln = _native.lib().chp_update_layout(n, *[_native.ptr(x) for x in a], _native.ptr(upd),
                                         _native.ptr(lev))
    assert ln >= 0
    return int(ln), upd, lev


def _forest():
    """A block-diagonal matrix of 12 grids: an elimination forest."""
    import scipy.sparse as sp
    return sp.block_diag(problems.batch_grids(12, 8, 30, seed=2), format="csc")


@pytest.mark.parametrize("make", [lambda: problems.grid3d(16, seed=1),
                                  lambda: problems.grid2d(64, seed=0),
                                  lambda: problems.band_arrow(3000, 8, 20, seed=4),
                                  _forest])
def test_update_slots_disjoint_while_live(make):
    A = make()
    symb = cp.symbolic(A, p=cp.amd.order)
    ln, upd, lev = _layout(symb)
    na = np.diff(symb.relptr).astype(np.int64)
    size = na * na
    par = np.asarray(symb.parent(), dtype=np.int64)
    end = lev[par]
    nonroot = par != np.arange(symb.Nsn)
    assert np.all(lev[par][nonroot] > lev[nonroot])
    assert np.all(upd + size <= ln)
    assert ln <= int(size.sum())
    for L in range(int(lev.max()) + 1):
        live = np.nonzero((lev <= L) & (end >= L) & (size > 0))[0]
        o = np.argsort(upd[live], kind="stable")
        s, e = upd[live][o], (upd + size)[live][o]
        assert np.all(e[:-1] <= s[1:]), "overlapping live slots at level %d" % L


def test_grid_arena_shrinks():
    A = problems.grid3d(24, seed=1)
    symb = cp.symbolic(A, p=cp.amd.order)
    ln, _, _ = _layout(symb)
    full = float((np.diff(symb.relptr).astype(float) ** 2).sum())
    print("arena %.1f M doubles of sum(na^2) %.1f M" % (ln / 1e6, full / 1e6))
    assert ln < 0.7 * full
