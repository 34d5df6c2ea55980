"""The flow-net sampling pattern (``hpc_fracture.sunflower_spiral``, ``hpc_fracture.py:21-48``):
properties that fix it without restating its formula."""

import numpy as np

from andfn_amd.hpc_fracture import sunflower_spiral


def test_sunflower_pattern():
    n_in, n_bnd = 500, 40
    z = sunflower_spiral(n_in, n_bnd)
    assert z.shape == (n_in + n_bnd,) and z.dtype == np.complex128
    inner, rim = z[:n_in], z[n_in:]
    r = np.abs(inner)
    # equal-area rings: the k-th point (from 0) sits where the disk area inside it is (k + 1/2)/n
    assert np.allclose(np.sort(r) ** 2, (np.arange(n_in) + 0.5) / n_in, rtol=0, atol=1e-14)
    assert np.all(np.diff(r) > 0)
    # consecutive points turn by the golden angle 2 pi / phi^2 (mod 2 pi, either way round)
    turn = np.angle(inner[1:] / inner[:-1])
    golden = 2 * np.pi * (1 - 2 / (1 + np.sqrt(5)))
    assert np.allclose(np.abs(turn), golden, rtol=0, atol=1e-9)
    # the rim: unit circle, first and last point at angle 0 and 2 pi, equal steps between
    assert np.allclose(np.abs(rim), 1.0) and np.isclose(rim[0], 1.0) and np.isclose(rim[-1], 1.0)
    steps = np.angle(rim[1:] / rim[:-1])
    assert np.allclose(steps, 2 * np.pi / (n_bnd - 1))
