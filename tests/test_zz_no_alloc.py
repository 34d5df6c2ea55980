"""The allocation-free heads entry point (andfn_mi_state_heads_noalloc) against andfn_mi_state_heads.

Both evaluate h = Re Omega_f(z) / t_f at the handle's solved state with the same kernel and the
same grouping of points into tiles, so the heads must be identical to the bit (NaN inside a well
included).  Two calls in a row, the second with fewer points, check that the shared scratch
carries nothing over; a call beyond the scratch's 65,536 points is refused.
"""

import numpy as np
import pytest

from andfn_amd import _lib
from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork

pytestmark = pytest.mark.gpu


def _noalloc(net, frac, z):
    frac = np.ascontiguousarray(frac, dtype=np.int64)
    zz = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.full(len(zz), -1.0)
    _lib.check(net.lib.andfn_mi_state_heads_noalloc(net.h, len(zz), _lib.ptr(frac), _lib.ptr(zz), _lib.ptr(out)))
    return out


def _points(fa, n, seed):
    rng = np.random.default_rng(seed)
    frac = rng.integers(0, len(fa), n)
    z = fa["radius"][frac] * rng.uniform(0.0, 0.9, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return frac, z


def test_noalloc_heads_equal_state_heads():
    print("no-alloc heads: dfn40w", flush=True)
    d = build_network(workload("dfn40w", constants=dict(MAX_ITERATIONS=200)))
    d.consolidate_dfn()
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        info = net.run(head_tol=1e-6)
        assert info["converged"]
        for n, seed in ((20000, 1), (333, 2)):
            frac, z = _points(fa, n, seed)
            # a well centre: the head there is NaN on both paths
            w = int(np.nonzero(ea["_type"] == 2)[0][0])
            frac[0], z[0] = ea["frac0"][w], ea["center"][w]
            ref = net.state_heads(frac, z)
            got = _noalloc(net, frac, z)
            print(f"{n} points: {int(np.sum(np.isnan(ref)))} NaN", flush=True)
            assert np.isnan(ref[0])
            assert np.array_equal(got, ref, equal_nan=True)
        with pytest.raises(_lib.CapacityError):
            _noalloc(net, np.zeros(65537, dtype=np.int64), np.zeros(65537, dtype=np.complex128))
    finally:
        net.close()
