"""Connectivity of the fracture graph in the library (host C++, andfn_mi_components) against
SciPy's connected components, the labelling the host builder uses (andfn/dfn.py:1674-1695 keeps
the fractures connected to a boundary; the builder keeps the largest such component).

Runs on the CPU: the call is host code of the C ABI.  Random graphs with isolated nodes,
repeated edges and self loops; the labels must be identical (components numbered in the order
of their smallest node), which fixes which component wins a tie in size.
"""

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from andfn_amd import construction


@pytest.mark.parametrize("seed", range(12))
def test_components_equal_scipy(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 4000))
    m = int(rng.integers(0, 2 * n))
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    lab = construction.components(n, a, b)
    ncomp, ref = connected_components(coo_matrix((np.ones(m), (a, b)), shape=(n, n)), directed=False)
    assert np.array_equal(lab, ref)
    assert lab.max() + 1 == ncomp


def test_components_rejects_bad_edges():
    from andfn_amd._lib import AndfnMiError

    with pytest.raises(AndfnMiError):
        construction.components(3, [0, 5], [1, 2])
