"""DFN construction through the C ABI: candidate pairs and chords on the GPU, connectivity in C++.

The device counterparts of the host builder in ``geometry.py``, with the same results
(``include/andfn_mi.h``, "DFN construction"):

* ``candidate_pairs``: every pair i < k whose bounding spheres meet, in (i, k) order, the pairs
  AnDFN's loop over fracture pairs can find intersecting (``andfn/dfn.py:1437-1455``);
* ``chords``: the chord two disks share in both planes (``andfn/geometry_functions.py:302-356``),
  shortened by SE_FACTOR (``:974-993``) and optionally clipped to rho times the radii;
* ``components``: connected components of the fracture graph (``andfn/dfn.py:1674-1695``).
"""

import ctypes

import numpy as np

from . import _lib


class DiskSet:
    """Disks as contiguous float64 arrays, kept alive with the ctypes struct that points at them."""

    def __init__(self, center, normal, x_vector, y_vector, radius):
        self.arrays = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
                       for a in (center, normal, x_vector, y_vector)]
        self.radius = np.ascontiguousarray(np.asarray(radius, dtype=np.float64).reshape(-1))
        n = len(self.radius)
        if any(len(a) != n for a in self.arrays):
            raise ValueError("disk arrays of different lengths")
        c, nv, xv, yv = self.arrays
        self.struct = _lib.Disks(n, c.ctypes.data, nv.ctypes.data, xv.ctypes.data, yv.ctypes.data,
                                 self.radius.ctypes.data)

    def __len__(self):
        return len(self.radius)


def candidate_pairs(disks, tol=1e-8, device=0):
    """(P, 2) int64 pairs (i, k), i < k, with |c_i - c_k| <= (r_i + tol) + (r_k + tol), in (i, k) order."""
    lib = _lib.load()
    count = ctypes.c_int64(0)
    cap = 8 * len(disks) + 1024
    while True:
        pi = np.empty(cap, dtype=np.int64)
        pk = np.empty(cap, dtype=np.int64)
        rc = lib.andfn_mi_candidate_pairs(device, ctypes.byref(disks.struct), tol, cap, _lib.ptr(pi), _lib.ptr(pk),
                                          ctypes.byref(count))
        if rc == _lib.ECAPACITY and count.value > cap:
            cap = count.value
            continue
        _lib.check(rc)
        n = count.value
        return np.stack([pi[:n], pk[:n]], axis=1)


def chords(ia, ib, a, b, shorten=3, se_factor=1.0, clip=0, rho=1.0, device=0):
    """Chords of (a[ia[p]], b[ib[p]]): endpoints (P, 2) complex in a's and in b's plane, and ok (P,) bool.

    shorten: bit 0 shrinks the chord in a's plane, bit 1 in b's plane, about its midpoint by
    se_factor; clip 1 keeps the part inside rho r of both disks, 2 inside rho r of disk b.
    """
    lib = _lib.load()
    ia = np.ascontiguousarray(ia, dtype=np.int64)
    ib = np.ascontiguousarray(ib, dtype=np.int64)
    n = len(ia)
    ea = np.empty((n, 2), dtype=np.complex128)
    eb = np.empty((n, 2), dtype=np.complex128)
    ok = np.empty(n, dtype=np.int8)
    _lib.check(lib.andfn_mi_disk_chords(device, n, _lib.ptr(ia), _lib.ptr(ib), ctypes.byref(a.struct),
                                        ctypes.byref(b.struct), int(shorten), float(se_factor), int(clip),
                                        float(rho), _lib.ptr(ea), _lib.ptr(eb), _lib.ptr(ok)))
    return ea, eb, ok.astype(bool)


def components(n, a, b):
    """Component label of every node of the graph (n nodes, edges (a[j], b[j])), numbered in the
    order of each component's smallest node."""
    lib = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    lab = np.empty(n, dtype=np.int64)
    _lib.check(lib.andfn_mi_components(int(n), len(a), _lib.ptr(a), _lib.ptr(b), _lib.ptr(lab)))
    return lab
