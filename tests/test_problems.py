"""Structure of the synthetic patterns the GPU tests and the benchmark use:
the chain patterns are chordal in natural order, and the caterpillar's spine
supernodes each sit alone on an elimination-tree level with two children."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import chompack_amd as cp
from chompack_amd import problems


def _heights(symb):
    par = np.asarray(symb.parent())
    h = np.zeros(symb.Nsn, dtype=int)
    for k in range(symb.Nsn):  # supernodes are postordered: children first
        if par[k] != k:
            h[par[k]] = max(h[par[k]], h[k] + 1)
    return h


def test_caterpillar_is_a_chain_with_pendant_children():
    spine, bw, arrows = 60, 4, 5
    symb = cp.symbolic(problems.caterpillar(spine, bw, arrows, seed=3), p=None)
    assert symb.n == 2 * spine + arrows
    assert symb.fill == (0, 0)  # natural order is a perfect elimination order
    h = _heights(symb)
    counts = np.bincount(h)
    assert counts[0] == spine  # the leaves
    assert np.all(counts[1:] == 1)  # one front per level above the leaves
    nch = np.diff(np.asarray(symb.chptr))
    assert np.sum(nch == 2) >= spine - bw - arrows  # spine fronts: leaf + previous spine front
    nj = np.diff(np.asarray(symb.snptr)) + np.diff(np.asarray(symb.relptr))
    assert nj.max() <= 64


def test_band_arrow_is_chordal_chain():
    symb = cp.symbolic(problems.band_arrow(400, 8, 10, seed=4), p=None)
    assert symb.fill == (0, 0)
    h = _heights(symb)
    assert np.all(np.bincount(h) == 1)  # a chain: one supernode per level
