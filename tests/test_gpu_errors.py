"""Error behaviour of the GPU solve at the reference's boundary (hpc_solve.py:55-257): an exactly
singular discharge matrix raises ``RuntimeError`` as SuperLU does (the oracle test of the same
records pins that), a coefficient count beyond the record capacity is refused before any work,
and a solve that reaches MAX_ITERATIONS returns its state with a warning."""

import logging

import numpy as np
import pytest

import errcases
import golden_io
from andfn_amd._lib import AndfnMiError
from andfn_amd.hpc_solve import last_solve_info, solve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("factor", ["sparse", "tournament", "rocsolver"])
def test_exactly_singular_discharge_matrix_raises(factor):
    fa, ea, consts, K = errcases.with_isolated_fracture()
    with pytest.raises(RuntimeError, match="singular"):
        solve(fa, ea, K, consts, factor=factor)


def test_coefficient_capacity_is_checked_before_solving():
    # capacity 30 (discharge_int 50 needs at least 25): MAX_NCOEF 30 + COEF_INCREASE 5 overflows
    fa, ea, consts, K = golden_io.records(golden_io.load("template"), max_ncoef=30)
    consts = consts.copy()
    consts["MAX_NCOEF"] = 30
    consts["COEF_INCREASE"] = 5
    coef0 = ea["coef"].copy()
    with pytest.raises(AndfnMiError) as err:
        solve(fa, ea, K, consts)
    assert err.value.code == -3
    assert np.array_equal(ea["coef"], coef0)


def test_max_iterations_returns_state_with_a_warning(caplog):
    fa, ea, consts, K = golden_io.records(golden_io.load("template"))
    consts = consts.copy()
    consts["MAX_ITERATIONS"] = 3
    with caplog.at_level(logging.WARNING, logger="andfn"):
        solve(fa, ea, K, consts)
    info = last_solve_info()
    assert not info["converged"] and info["nit"] == 3
    assert any("MAX_ITERATIONS" in r.getMessage() for r in caplog.records)
    assert np.all(np.isfinite(ea["q"])) and np.any(ea["q"] != 0.0)
