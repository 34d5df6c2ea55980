"""Warm start from a checkpoint (andfn/dfn.py:404-609; consolidate_dfn copies the saved
coefficients and discharges into the records, andfn/dfn.py:727,734).

A network solved to a head residual of 1e-9 m is saved, loaded into a new DFN and solved again
with the same tolerance: the loaded state already meets it, so the second solve stops after its
first iteration, and its coefficients, discharges and constants stay within 1e-7 of the first
solve's (relative to the largest magnitude: that one damped iteration moves a state at head
residual 1e-9 m by ~5e-9).  A cold solve of the same network takes ~100 iterations.
"""

import numpy as np
import pytest

from andfn_amd.dfn import DFN
from andfn_amd.generator import build_network, workload

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["cfg1", "dfn40w"])
def test_warm_start_from_checkpoint(tmp_path, name):
    d = build_network(workload(name, constants=dict(MAX_ITERATIONS=3000, MAX_ERROR=0.0)))
    d.solve(head_tol=1e-9)
    cold = d.last_solve_info
    assert cold["converged"] and cold["nit"] > 5
    fn = d.save_dfn(str(tmp_path / f"{name}.npz"))
    e = DFN("warm").load_dfn(fn)
    e.solve(head_tol=1e-9)
    warm = e.last_solve_info
    print(name, "cold nit", cold["nit"], "warm nit", warm["nit"], "residual", warm["head_residual"])
    assert warm["converged"] and warm["nit"] <= 2
    ea, eb = d.elements_struc_array_hpc, e.elements_struc_array_hpc
    assert np.array_equal(ea["ncoef"], eb["ncoef"])
    assert _rel(eb["coef"], ea["coef"]) <= 1e-7 and _rel(eb["q"], ea["q"]) <= 1e-7
    assert _rel(e.fractures_struc_array_hpc["constant"], d.fractures_struc_array_hpc["constant"]) <= 1e-7
