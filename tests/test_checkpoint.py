"""Checkpoint of a network with its solution (andfn/dfn.py:404-609): save_dfn / load_dfn.

A network with every element type and a solution state (coefficients, counts, discharges,
errors, fracture constants) is saved and loaded into a new DFN; the consolidated records must be
identical field by field, so a solve of the loaded network starts from the saved state (the
warm start).  The .npz container is written and read here; the .h5 container needs h5py, which
this environment lacks (that path follows AnDFN's layout, one dataset per record and field, and
stays untested here).
"""

import numpy as np
import pytest

from andfn_amd.dfn import DFN, ImpermeableCircle, ImpermeableLine
from andfn_amd.generator import build_network, workload


def _network_with_state(seed):
    d = build_network(workload("dfn40w"))
    f = d.fractures[0]
    ImpermeableCircle("impc", 0.3, 0.1 * f.radius + 0.05j * f.radius, f, 6, 12)
    ImpermeableLine("impl", np.array([-0.2 + 0.4j, 0.1 + 0.45j]) * f.radius, f, 6, 12)
    d.get_elements()
    rng = np.random.default_rng(seed)
    for e in d.elements:
        e.ncoef = int(rng.integers(1, 12))
        e.nint = 2 * e.ncoef
        e.coef = rng.standard_normal(e.ncoef) + 1j * rng.standard_normal(e.ncoef)
        e.q = float(rng.standard_normal())
        e.error = float(rng.random())
    for f in d.fractures:
        f.constant = float(rng.standard_normal())
        f.aperture = float(rng.random()) if rng.random() < 0.5 else None
    d.set_kwargs(MAX_ITERATIONS=77, DAMPING=0.4)
    d.consolidate_dfn()
    return d


@pytest.mark.parametrize("seed", [0, 1])
def test_save_load_reproduces_records(tmp_path, seed):
    d = _network_with_state(seed)
    fn = d.save_dfn(str(tmp_path / "net.npz"))
    e = DFN("loaded").load_dfn(fn)
    for name in d.elements_struc_array_hpc.dtype.names:
        assert np.array_equal(d.elements_struc_array_hpc[name], e.elements_struc_array_hpc[name]), name
    for name in d.fractures_struc_array_hpc.dtype.names:
        assert np.array_equal(d.fractures_struc_array_hpc[name], e.fractures_struc_array_hpc[name]), name
    assert [x.label for x in d.elements] == [x.label for x in e.elements]
    assert [x.aperture for x in d.fractures] == [x.aperture for x in e.fractures]
    assert e.discharge_int == d.discharge_int
    assert int(e.constants["MAX_ITERATIONS"]) == 77 and float(e.constants["DAMPING"]) == 0.4


def test_save_refuses_to_overwrite_and_checks_extension(tmp_path):
    d = _network_with_state(0)
    fn = str(tmp_path / "net.npz")
    assert d.save_dfn(fn) == fn
    assert d.save_dfn(fn) is None  # exists: AnDFN warns and returns
    assert d.save_dfn(fn, overwrite=True) == fn
    with pytest.raises(ValueError):
        d.save_dfn(str(tmp_path / "net.txt"))
    with pytest.raises(FileNotFoundError):
        DFN("x").load_dfn(str(tmp_path / "missing.npz"))
