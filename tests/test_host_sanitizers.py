"""The host C++ of the symbolic layer (amd.cpp, symbolic.cpp) under
AddressSanitizer and UndefinedBehaviorSanitizer (CPU, no GPU).

The two sources are built with g++ into a host-only shared library with
-fsanitize=address,undefined (halting on the first report), loaded through
the package's own ctypes declarations in a child process that preloads the
ASan runtime, and run over the symbolic pipeline (AMD ordering, elimination
tree, postorder, column counts, supernodes, amalgamation, clique embedding,
relative indices), maximum cardinality search, maxchord and lmerge on a set
of patterns: grids, the reference test matrix, chains, arrows, random
patterns, a matrix with empty columns and the 1 x 1 case.  The child's
arrays must equal the ones the regular library gives (a digest per case),
and any sanitizer report fails the test.
"""
import hashlib
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "chompack_amd", "csrc")

CHILD = r'''
import ctypes, hashlib, sys
import numpy as np, scipy.sparse as sp
sys.path.insert(0, sys.argv[1])
import chompack_amd._native as nat
if len(sys.argv) > 2:
    L = ctypes.CDLL(sys.argv[2])
    nat._declare(L)
    nat._lib = L
import chompack_amd as cp
from chompack_amd import problems, misc

def rand_pattern(n, dens, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=dens, random_state=rng, format="coo")
    A = sp.tril(A + A.T + sp.eye(n) * n, format="csc")
    return A

cases = [
    ("grid2d_30", problems.grid2d(30, seed=0)),
    ("grid3d_9", problems.grid3d(9, seed=1)),
    ("ref31", problems.reference_test_matrix()),
    ("two_cliques", problems.two_cliques()),
    ("caterpillar", problems.caterpillar()),
    ("band_arrow", problems.band_arrow(400, 5, 7, seed=2)),
    ("rand_200", rand_pattern(200, 0.02, 3)),
    ("rand_500", rand_pattern(500, 0.005, 4)),
    ("diag_only", sp.eye(50, format="csc")),
    ("one", sp.csc_matrix(np.array([[2.0]]))),
]
for name, A in cases:
    h = hashlib.sha1()
    for p in (None, cp.amd.order):
        s = cp.symbolic(A, p=p)
        for arr in (s.snpost, s.snptr, s.relptr, s.relidx, s.chptr, s.chidx, s.blkptr, s.p):
            h.update(np.ascontiguousarray(arr, dtype=np.int64).tobytes())
    S = misc.symmetrize(A)
    h.update(np.asarray(cp.maxcardsearch(S), dtype=np.int64).tobytes())
    Ac, pc = cp.maxchord(S)
    h.update(np.asarray(pc, dtype=np.int64).tobytes())
    h.update(np.asarray(sp.csc_matrix(Ac).indices, dtype=np.int64).tobytes())
    a = np.arange(0, 40, 3, dtype=np.int64)
    b = np.arange(1, 50, 4, dtype=np.int64)
    left = np.zeros(len(a) + len(b), dtype=np.int64)
    left[:len(a)] = a
    k = misc.lmerge(left, b, 0, 0, len(a), len(b))
    h.update(left[:k].tobytes())
    print(name, h.hexdigest(), flush=True)
'''


def _asan_runtime():
    gxx = shutil.which("g++")
    if not gxx:
        return None, None
    lib = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return gxx, (lib if os.path.isabs(lib) and os.path.exists(lib) else None)


def test_symbolic_host_code_under_sanitizers(tmp_path):
    gxx, asan = _asan_runtime()
    if not gxx or not asan:
        pytest.skip("g++ with the ASan runtime is not available")
    so = str(tmp_path / "libhost_san.so")
    build = [gxx, "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             os.path.join(CSRC, "symbolic.cpp"), os.path.join(CSRC, "amd.cpp"), "-o", so]
    r = subprocess.run(build, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    env = dict(os.environ)
    plain = subprocess.run([sys.executable, str(script), ROOT], capture_output=True, text=True, env=env,
                           timeout=600)
    assert plain.returncode == 0, plain.stderr[-3000:]
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    san = subprocess.run([sys.executable, str(script), ROOT, so], capture_output=True, text=True, env=env,
                         timeout=900)
    assert san.returncode == 0, san.stderr[-3000:]
    assert "runtime error" not in san.stderr and "AddressSanitizer" not in san.stderr, san.stderr[-3000:]
    assert san.stdout.split() == plain.stdout.split() and plain.stdout.strip()
