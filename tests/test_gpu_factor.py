"""The library's multifrontal sparse LU (the default) and its dense blocked LU with tournament
pivoting, against rocSOLVER's dense getrf.

All three factor the same equilibrated discharge matrix with different orderings and row
pivoting, so the per-iteration discharge solves agree to rounding and so do the iterates.  Networks of the 200-disk
law at constant density: 40 fractures (one panel, n < 64), 1,500 fractures (3,470 unknowns: two
tournament levels) and 4,000 fractures (8,884 unknowns: three levels, 139 panels).  Fixed
iteration count (the change-based stop off), so both runs take the same iterations.  A batch of
three realisations (three discharge blocks factored in one batch) is compared the same way.
Tolerance: 1e-10 relative to the largest magnitude of each quantity; the head residual of the
final iterate (still hundreds of metres after 40 iterations) agrees to 1e-10 relative.
"""

import numpy as np
import pytest

from andfn_amd import _lib
from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork

pytestmark = pytest.mark.gpu
CONSTS = dict(MAX_ERROR=0.0, MAX_ITERATIONS=40)


def _rel(a, b):
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def _network(n):
    if n == 40:
        # coefficient growth off, as in the other networks: the comparison is of the factorisation,
        # not of a growth trajectory (whose threshold tests amplify rounding-level differences)
        spec = workload("dfn40", constants=dict(CONSTS, MAX_NCOEF=10))
    else:
        spec = workload("cfg1", n_fractures=n, cube=100.0 * (n / 200.0) ** (1.0 / 3.0), constants=CONSTS)
    d = build_network(spec)
    d.consolidate_dfn()
    return d


def _run(d, factor, precision="fp64", head_tol=0.0):
    fa = d.fractures_struc_array_hpc.copy()
    ea = d.elements_struc_array_hpc.copy()
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        info = net.run(precision, head_tol=head_tol, factor=factor)
        net.download()
    finally:
        net.close()
    return fa, ea, info


@pytest.mark.parametrize("factor", ["tournament", "sparse"])
@pytest.mark.parametrize("n", [40, 1500, 4000])
def test_lu_matches_rocsolver(n, factor):
    print(f"network of {n} fractures, {factor}", flush=True)
    d = _network(n)
    fr, er, ir = _run(d, "rocsolver")
    ft, et, it = _run(d, factor)
    print(f"unknowns {it['n_unknowns']}, factor s: rocsolver {ir['t_factor']:.3f} {factor} {it['t_factor']:.3f}, "
          f"stored factor entries {it['nnz_factors']}")
    assert ir["nit"] == it["nit"] == CONSTS["MAX_ITERATIONS"]
    assert _rel(et["coef"], er["coef"]) < 1e-10
    assert _rel(et["q"], er["q"]) < 1e-10
    assert _rel(ft["constant"], fr["constant"]) < 1e-10
    assert abs(it["head_residual"] - ir["head_residual"]) <= 1e-10 * max(1.0, ir["head_residual"])


@pytest.mark.parametrize("factor", ["tournament", "sparse"])
def test_mixed_form_reaches_head_tolerance(factor):
    d = _network(1500)
    c = np.array(d.constants, copy=True)
    c["MAX_ITERATIONS"] = 400
    d.constants = c
    _, _, info = _run(d, factor, precision="mixed", head_tol=1e-6)
    assert info["converged"] and info["head_residual"] <= 1e-6


@pytest.mark.parametrize("factor", ["tournament", "sparse"])
def test_batch_of_realisations_matches_rocsolver(factor):
    # three independent realisations: one discharge block each, factored in one batch
    from andfn_amd.batch import build_batch, realisation_specs

    specs = realisation_specs(3, seed=7, nmin=150, nmax=400)
    _, fa, ea, off, K, c = build_batch(specs, dict(MAX_ERROR=0.0, MAX_NCOEF=10, MAX_ITERATIONS=40))
    res = {}
    for fct in ("rocsolver", factor):
        f1, e1 = fa.copy(), ea.copy()
        net = DeviceNetwork(f1, e1, K, c)
        try:
            info = net.run("fp64", factor=fct)
            net.download()
        finally:
            net.close()
        res[fct] = (f1, e1, info)
    (fr, er, ir), (ft, et, it) = res["rocsolver"], res[factor]
    assert ir["nit"] == it["nit"] == 40
    assert _rel(et["coef"], er["coef"]) < 1e-10
    assert _rel(et["q"], er["q"]) < 1e-10
    assert _rel(ft["constant"], fr["constant"]) < 1e-10


def test_unknown_factor_is_rejected():
    with pytest.raises(ValueError):
        _run(_network(40), "qr")
    assert _lib.FACTORS == {"sparse": 0, "rocsolver": 1, "tournament": 2}
