"""The network builder against the inputs AnDFN itself consolidated for the same fractures.

For the synthetic golden cases, tests/golden/make_golden.py rebuilt the network with AnDFN's
own classes (Fracture, get_fracture_intersections, set_constant_head_boundary, Well) from the
fracture list this package generates; the element records must agree: same element order,
types, fracture links and per-fracture element lists, and endpoints/heads/potentials to
rounding (the plane-intersection formulas amplify last-bit differences of dot products through
the slope form of the line-circle intersection, hence a 1e-9 absolute tolerance in metres; the
largest difference seen is 2.7e-10 m on the 200-fracture case, ~1e-11 of the fracture radii).
"""

import numpy as np
import pytest

import golden_io
from andfn_amd.generator import build_network, workload

SYN = [c for c in golden_io.CASES if c.startswith("dfn")]


@pytest.mark.parametrize("case", SYN)
def test_builder_matches_andfn_records(case):
    g = golden_io.load(case)
    d = build_network(workload(case))
    d.consolidate_dfn()
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    assert len(ea) == len(g["in_e__type"]) and len(fa) == len(g["in_f_t"])
    for k in ("_type", "frac0", "frac1", "ncoef", "nint"):
        assert np.array_equal(ea[k], g["in_e_" + k]), k
    assert np.array_equal(fa["nelements"], g["in_f_nelements"])
    for f in range(len(fa)):
        n = fa["nelements"][f]
        assert np.array_equal(fa["elements"][f, :n], g["in_f_elements"][f, :n])
    for k in ("endpoints0", "endpoints1", "center"):
        assert np.max(np.abs(ea[k] - g["in_e_" + k])) < 1e-9, k
    for k in ("radius", "head", "phi"):
        assert np.allclose(ea[k], g["in_e_" + k], rtol=1e-12, atol=1e-14), k
    for k in ("t", "radius", "center", "normal", "x_vector", "y_vector"):
        assert np.allclose(fa[k], g["in_f_" + k], rtol=1e-13, atol=1e-13), k
