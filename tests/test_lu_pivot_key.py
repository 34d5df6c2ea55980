"""The tournament LU's pivot key (andfn_amd/csrc/lu_pivot_key.hpp), built for the host with g++.

One 64-bit max over the keys of a row group must select the row of largest |a| (ties within
2^-42 relative go to the lowest thread), for every magnitude FP64 holds down to ~5e-321
(1024 subnormal steps, the bits the thread index takes): subnormals and values far below the
smallest float (a float key rounds everything below ~1e-45 to 0 and ties it with an
exact zero), zeros (an active row's key is never 0, the mark of a row that takes no part) and NaN
(ranked above every number).  The expected winner is found by plain comparison of |a|.
"""

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

HDR = Path(__file__).resolve().parents[1] / "andfn_amd" / "csrc" / "lu_pivot_key.hpp"
SRC = r'''
#include "lu_pivot_key.hpp"
extern "C" unsigned long long key(double a, int t) { return aem::lu_pivot_key(a, t); }
extern "C" int thread_of(unsigned long long k) { return aem::lu_pivot_thread(k); }
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("pivkey")
    (d / "k.cpp").write_text(SRC)
    so = d / "libk.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", str(HDR.parent), str(d / "k.cpp"), "-o",
                    str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.key.restype = ctypes.c_ulonglong
    L.key.argtypes = [ctypes.c_double, ctypes.c_int]
    L.thread_of.restype = ctypes.c_int
    L.thread_of.argtypes = [ctypes.c_ulonglong]
    return L


def _winner(L, col):
    keys = [L.key(float(a), t) for t, a in enumerate(col)]
    assert all(k != 0 for k in keys)
    best = max(keys)
    return L.thread_of(best)


def test_thread_index_round_trips(lib):
    for t in (0, 1, 7, 255, 511):
        for a in (0.0, -0.0, 1e-320, 3.5, -2.0e300):
            assert lib.thread_of(lib.key(a, t)) == t


def test_largest_magnitude_wins_down_to_subnormals(lib):
    rng = np.random.default_rng(0)
    for scale in (1.0, 1e-40, 1e-100, 1e-300, 5e-324 * 1e6):
        col = np.zeros(512)
        col[rng.choice(512, 40, replace=False)] = rng.standard_normal(40) * scale
        w = _winner(lib, col)
        assert abs(col[w]) == np.max(np.abs(col)), scale
    # a tiny nonzero beats exact zeros wherever it sits
    for pos in (0, 100, 511):
        col = np.zeros(512)
        col[pos] = -1e-46
        assert _winner(lib, col) == pos
        col[pos] = 1e-318  # subnormal; only magnitudes below 2^10 subnormal steps (~5e-321) rank as 0
        assert _winner(lib, col) == pos


def test_ties_go_to_the_lowest_thread_and_nan_ranks_first(lib):
    col = np.zeros(512)
    col[[30, 90, 400]] = [2.0, -2.0, 2.0]
    assert _winner(lib, col) == 30
    col[200] = np.inf
    assert _winner(lib, col) == 200
    col[300] = np.nan
    assert _winner(lib, col) == 300
    # magnitudes 2^-43 apart still rank by magnitude, not by thread
    col = np.zeros(512)
    col[10] = 1.0
    col[20] = 1.0 + 2.0 ** -40
    assert _winner(lib, col) == 20
