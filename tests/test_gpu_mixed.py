"""Mixed-precision form against the FP64 form on the GPU.

The mixed form factors the equilibrated discharge matrix in FP32 and refines against the FP64
residual inside the fixed-point loop, and evaluates far-field line sources (|Z| >= 4 for the
whole wavefront) and the coefficient fits in FP32 until the head residual is near the
tolerance; then it finishes in FP64.  Both forms stop at a head residual of 1e-6 m, so they
are two approximations of one fixed point that differ by an amount of the order of that
residual.

Coefficient growth reacts to rounding-level differences of the error estimates, so the states
are compared where the coefficient counts agree: with growth off, and with the mixed run
replaying the FP64 run's counts.  Tolerances (relative to the largest magnitude of each quantity
in the FP64 result): 1e-6 for coefficients, discharges and fracture constants; heads at 100
observation points within 1e-5 m.
"""

import numpy as np
import pytest

from andfn_amd import _lib
from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork, get_heads, last_solve_info, solve

pytestmark = pytest.mark.gpu
HEAD_TOL = 1e-6
REL_TOL = 1e-6
HEAD_ABS_TOL = 1e-5


def _rel(a, b):
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def _observation_points(fa, n, seed):
    rng = np.random.default_rng(seed)
    frac = rng.integers(0, len(fa), n)
    rad = fa["radius"][frac] * rng.uniform(0.1, 0.8, n)
    z = rad * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return frac, z


def _pair(name, **consts):
    # AnDFN's example networks keep its MAX_ITERATIONS of 50; the head-residual stop needs more
    consts.setdefault("MAX_ITERATIONS", 500)
    d = build_network(workload(name, constants=consts))
    d.consolidate_dfn()
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    return fa, ea, fa.copy(), ea.copy(), d.discharge_int, d.constants


def _compare(fa, ea, fm, em, i64, imx):
    assert i64["head_residual"] <= HEAD_TOL and imx["head_residual"] <= HEAD_TOL, (i64, imx)
    assert imx["nit_fp32"] > 0, imx
    assert np.array_equal(em["ncoef"], ea["ncoef"])
    assert _rel(em["coef"], ea["coef"]) < REL_TOL
    assert _rel(em["q"], ea["q"]) < REL_TOL
    assert _rel(fm["constant"], fa["constant"]) < REL_TOL
    frac, z = _observation_points(fa, 100, 7)
    h64 = get_heads(fa, ea, frac, z)
    hmx = get_heads(fm, em, frac, z)
    assert np.all(np.isfinite(h64))
    assert np.max(np.abs(hmx - h64)) < HEAD_ABS_TOL


@pytest.mark.parametrize("name", ["cfg1"])
def test_mixed_matches_fp64_fixed_counts(name):
    print(f"mixed vs fp64, fixed coefficient counts, {name}", flush=True)
    fa, ea, fm, em, K, consts = _pair(name, MAX_NCOEF=int(workload(name).ncoef))
    solve(fa, ea, K, consts, head_tol=HEAD_TOL)
    i64 = last_solve_info()
    solve(fm, em, K, consts, precision="mixed", head_tol=HEAD_TOL)
    _compare(fa, ea, fm, em, i64, last_solve_info())


@pytest.mark.parametrize("name,consts", [("cfg1", {"MAX_NCOEF": 20}), ("dfn40w", {})])
def test_mixed_matches_fp64_on_fp64_growth_trajectory(name, consts):
    # coefficient growth reacts to the FP32 error estimates, so the mixed run replays the FP64
    # run's coefficient counts to make the two states comparable element by element
    print(f"mixed vs fp64, growth replay, {name}", flush=True)
    fa, ea, fm, em, K, consts = _pair(name, **consts)
    solve(fa, ea, K, consts, head_tol=HEAD_TOL, record_growth=True)
    i64 = last_solve_info()
    solve(fm, em, K, consts, precision="mixed", head_tol=HEAD_TOL, growth_schedule=i64["ncoef_history"])
    _compare(fa, ea, fm, em, i64, last_solve_info())


def test_mixed_repeated_runs_on_resident_network():
    print("mixed repeated runs cfg1", flush=True)
    d = build_network(workload("cfg1"))
    d.consolidate_dfn()
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        infos = [net.run("mixed", head_tol=HEAD_TOL) for _ in range(2)]
        net.download()
    finally:
        net.close()
    assert infos[0]["nit"] == infos[1]["nit"] and infos[0]["nit_fp32"] == infos[1]["nit_fp32"]
    assert infos[1]["head_residual"] <= HEAD_TOL


def test_mixed_requires_head_tolerance():
    print("mixed without head_tol", flush=True)
    d = build_network(workload("dfn40"))
    d.consolidate_dfn()
    with pytest.raises(_lib.AndfnMiError):
        solve(d.fractures_struc_array_hpc, d.elements_struc_array_hpc, d.discharge_int, d.constants, precision="mixed")
