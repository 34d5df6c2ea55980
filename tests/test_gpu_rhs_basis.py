"""The head right-hand side from its precomputed basis (rhs_basis_kernel / rhs_eval_kernel, the
default) against the sweep of the discharge points (the fallback, forced with
ANDFN_MI_RHS_SWEEP): the same map of the coefficients summed in another order, so the solves
must agree to rounding.  Growth trajectories are replayed from AnDFN's golden runs (so that a
rounding-level tie cannot change a growth decision), on cases with intersections, constant-head
lines, a well, and impermeable circles and lines."""

import os

import numpy as np
import pytest

import golden_io
from andfn_amd.hpc_solve import last_solve_info, solve

pytestmark = pytest.mark.gpu


def _solve(case, sweep):
    g = golden_io.load(case)
    fa, ea, consts, K = golden_io.records(g)
    if sweep:
        os.environ["ANDFN_MI_RHS_SWEEP"] = "1"
    try:
        solve(fa, ea, K, consts, growth_schedule=g.get("out_ncoef_history"))
    finally:
        os.environ.pop("ANDFN_MI_RHS_SWEEP", None)
    return fa, ea, last_solve_info()


@pytest.mark.parametrize("case", ["template_t", "template_imp", "dfn40w"])
def test_basis_matches_sweep(case):
    print(f"rhs basis vs sweep: {case}", flush=True)
    fa0, ea0, i0 = _solve(case, sweep=False)
    fa1, ea1, i1 = _solve(case, sweep=True)
    assert i0["rhs_points_per_lane"] == 0 and i1["rhs_points_per_lane"] > 0
    assert i0["nit"] == i1["nit"]
    assert np.array_equal(ea0["ncoef"], ea1["ncoef"])

    def rel(a, b):
        s = np.max(np.abs(b))
        return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))

    d = dict(coef=rel(ea0["coef"], ea1["coef"]), q=rel(ea0["q"], ea1["q"]),
             constant=rel(fa0["constant"], fa1["constant"]))
    print(case, d)
    assert d["coef"] < 1e-10 and d["q"] < 1e-10 and d["constant"] < 1e-10, d
