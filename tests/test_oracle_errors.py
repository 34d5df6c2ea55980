"""The reference's error behaviour at the solve boundary, restated by the oracle."""

import pytest

import errcases
from oracle import oracle


def test_exactly_singular_discharge_matrix_raises():
    fa, ea, consts, K = errcases.with_isolated_fracture()
    with pytest.raises(RuntimeError, match="singular"):
        oracle.solve(fa, ea, K, consts)
