"""DFN construction on the GPU (andfn_mi_candidate_pairs, andfn_mi_disk_chords) against the host
builder, which tests/test_inputs.py pins to AnDFN's own records.

* candidate pairs: the same pairs in the same (i, k) order as the host's KD-tree search, on the
  20,000-disk law (the pairs AnDFN's loop over i < k tests, andfn/dfn.py:1437-1455);
* chords (andfn/geometry_functions.py:302-356, shortened by :974-993): the same accept decision
  for every candidate pair and endpoints equal to the last bit (the kernels follow the host's
  arithmetic without multiply-add contraction); with the rim clipping, whose |z| is a hypot the
  two math libraries may round differently, the same decisions and endpoints within 1e-11 m
  (1e-13 of the largest radius);
* whole networks built with device=True and device=False: identical records for AnDFN's
  construction (no clipping) and the same structure with endpoints within 1e-12 m for the
  clipped benchmark workloads (1e-11 m), the cfg3 network included; AnDFN's own records for the golden
  cases through test_inputs.py's comparison.
"""

import time

import numpy as np
import pytest

import golden_io
from andfn_amd import construction
from andfn_amd import geometry as geo
from andfn_amd.generator import build_network, draw_fractures, workload

pytestmark = pytest.mark.gpu


def _disks(name):
    spec = workload(name)
    c, nr, r, _ = draw_fractures(spec)
    n, x, y = geo.fracture_basis(nr)
    return c, n, x, y, r


@pytest.mark.parametrize("name", ["cfg2", "cfg3"])
def test_candidate_pairs_equal_host_search(name):
    c, n, x, y, r = _disks(name)
    ds = construction.DiskSet(c, n, x, y, r)
    t0 = time.perf_counter()
    pd = construction.candidate_pairs(ds, 1e-8)
    t1 = time.perf_counter()
    ph = geo.candidate_pairs(c, r + 1e-8)
    t2 = time.perf_counter()
    print(f"{name}: {len(pd)} candidate pairs, device {t1 - t0:.3f} s, host {t2 - t1:.3f} s")
    assert np.array_equal(pd, ph)


@pytest.mark.parametrize("rim", [0.0, 0.03])
def test_chords_equal_host_formulas(rim):
    c, n, x, y, r = _disks("cfg3")
    ds = construction.DiskSet(c, n, x, y, r)
    p = construction.candidate_pairs(ds, 1e-8)
    i, k = p[:, 0], p[:, 1]
    se = 0.9
    e0d, e1d, okd = construction.chords(i, k, ds, ds, 3, se, 1 if rim > 0 else 0, 1.0 - rim)
    e0, e1, ok = geo.fracture_intersection(c[i], n[i], x[i], y[i], r[i], c[k], n[k], x[k], y[k], r[k])
    e0 = geo.shorten_line(e0, se)
    e1 = geo.shorten_line(e1, se)
    if rim > 0:
        e0, e1, ok2 = geo.clip_to_disks(e0, r[i], e1, r[k], 1.0 - rim)
        ok = ok & ok2
    print(f"rim {rim}: {len(p)} pairs, {int(ok.sum())} chords")
    assert np.array_equal(okd, ok)
    if rim == 0.0:
        assert np.array_equal(e0d[ok], e0[ok]) and np.array_equal(e1d[ok], e1[ok])
    else:
        assert np.max(np.abs(e0d[ok] - e0[ok])) <= 1e-11 and np.max(np.abs(e1d[ok] - e1[ok])) <= 1e-11


def _records(name, device):
    t0 = time.perf_counter()
    d = build_network(workload(name), device=device)
    d.consolidate_dfn()
    return d.fractures_struc_array_hpc, d.elements_struc_array_hpc, time.perf_counter() - t0


@pytest.mark.parametrize("name", ["dfn40", "dfn40w", "dfn200", "dfn5000", "cfg1", "cfg2", "cfg3"])
def test_device_built_network_equals_host_built(name):
    fd, ed, td = _records(name, True)
    fh, eh, th = _records(name, False)
    print(f"{name}: {len(fd)} fractures, {len(ed)} elements; built in {td:.2f} s (device) / {th:.2f} s (host)")
    assert len(ed) == len(eh) and len(fd) == len(fh)
    for k in ("_type", "frac0", "frac1", "ncoef", "nint"):
        assert np.array_equal(ed[k], eh[k]), k
    assert np.array_equal(fd["nelements"], fh["nelements"]) and np.array_equal(fd["elements"], fh["elements"])
    exact = workload(name).rim_clearance == 0.0
    for k in ("endpoints0", "endpoints1", "center", "radius", "phi"):
        if exact:
            assert np.array_equal(ed[k], eh[k]), k
        else:
            assert np.max(np.abs(ed[k] - eh[k])) <= 1e-11, k
    for k in ("t", "radius", "center", "normal", "x_vector", "y_vector"):
        assert np.array_equal(fd[k], fh[k]), k


@pytest.mark.parametrize("case", [c for c in golden_io.CASES if c.startswith("dfn")])
def test_device_built_network_matches_andfn_records(case):
    g = golden_io.load(case)
    fa, ea, _ = _records(case, True)
    for k in ("_type", "frac0", "frac1", "ncoef", "nint"):
        assert np.array_equal(ea[k], g["in_e_" + k]), k
    for f in range(len(fa)):
        m = fa["nelements"][f]
        assert np.array_equal(fa["elements"][f, :m], g["in_f_elements"][f, :m])
    for k in ("endpoints0", "endpoints1", "center"):
        assert np.max(np.abs(ea[k] - g["in_e_" + k])) < 1e-9, k
