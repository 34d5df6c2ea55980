"""The 20,000-fracture workload at full size (17,237 fractures kept, 50 wells, 65,690 unknowns).

The oracle's full solve of this network does not fit a test (SciPy's SuperLU alone takes ~9
minutes on it), so the GPU's converged states are checked through what the oracle can recompute
in seconds, its own discharge matrix and right-hand side at the GPU's state:

* FP64 form, stopped at head residual 1e-6 m: the oracle's head residual of the returned
  coefficients, discharges and constants is within the tolerance and equals the GPU's to 1e-3 of
  it; every fracture's discharges sum to zero (continuity rows) to 1e-10 of the largest |q|;
* mixed form at the same tolerance: the oracle's head residual of its state is within the
  tolerance, and coefficients, discharges and constants agree with the FP64 form to 1e-6
  relative (both are approximations of one fixed point at residual 1e-6 m), heads at 100
  observation points to 1e-5 m.
"""

import numpy as np
import pytest

from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork
from oracle import oracle

pytestmark = pytest.mark.gpu
HEAD_TOL = 1e-6


@pytest.fixture(scope="module")
def cfg3():
    print("cfg3: building the network", flush=True)
    d = build_network(workload("cfg3"))
    d.consolidate_dfn()
    return d


def _obs(fa, n, seed=3):
    rng = np.random.default_rng(seed)
    of = rng.integers(0, len(fa), n)
    oz = fa["radius"][of] * rng.uniform(0.1, 0.7, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return of, oz


def _run(d, precision):
    fa = d.fractures_struc_array_hpc.copy()
    ea = d.elements_struc_array_hpc.copy()
    print(f"cfg3: {precision} solve", flush=True)
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        info = net.run(precision=precision, head_tol=HEAD_TOL)
        net.download()
        of, oz = _obs(fa, 100)
        heads = net.state_heads(of, oz)
    finally:
        net.close()
    print(f"cfg3 {precision}: nit {info['nit']}, head residual {info['head_residual']:.3e}, "
          f"factor {info['t_factor']:.3f}s, iterate {info['t_iterate']:.3f}s", flush=True)
    return fa, ea, info, heads


@pytest.fixture(scope="module")
def fp64(cfg3):
    return _run(cfg3, "fp64")


def _oracle_rows(d, fa, ea):
    """(|A y - b| of the head rows in metres, the continuity rows) of the oracle's discharge
    system at the state held by the records (hpc_solve.py:832-887 for A, 155-174 for b)."""
    O = oracle.OracleHandle(fa.copy(), ea.copy(), d.discharge_int)
    try:
        y = np.zeros(O.n)
        O.lib.oracle_old_discharges(O.h, oracle._ptr(y))
        A = O.discharge_matrix()
        b = O.head_rhs()
        nde = O.nde
    finally:
        O.close()
    r = A @ y - b
    de = np.nonzero(np.isin(ea["_type"], (0, 2, 3)))[0]
    scale = np.ones(nde)
    chl = ea["_type"][de] != 0
    scale[chl] = 1.0 / fa["t"][ea["frac0"][de[chl]]]
    return np.abs(r[:nde]) * scale, r[nde:]


def test_cfg3_fp64_head_residual_recomputed_by_the_oracle(cfg3, fp64):
    fa, ea, info, _ = fp64
    assert info["converged"] and info["head_residual"] <= HEAD_TOL
    assert info["n_unknowns"] == int(np.sum(np.isin(ea["_type"], (0, 2, 3)))) + len(fa)
    assert int(np.sum(ea["_type"] == 2)) == 50
    print("cfg3: oracle rows", flush=True)
    head, cont = _oracle_rows(cfg3, fa, ea)
    print(f"cfg3 fp64: GPU head residual {info['head_residual']:.6e}, oracle {head.max():.6e}; "
          f"continuity {np.max(np.abs(cont)):.3e} (max |q| {np.max(np.abs(ea['q'])):.3e})")
    assert head.max() <= HEAD_TOL * (1 + 1e-6)
    assert abs(head.max() - info["head_residual"]) <= 1e-3 * HEAD_TOL
    assert np.max(np.abs(cont)) <= 1e-10 * np.max(np.abs(ea["q"]))


def test_cfg3_mixed_against_fp64_and_the_oracle(cfg3, fp64):
    fa, ea, info, h64 = fp64
    fam, eam, infom, hm = _run(cfg3, "mixed")
    assert infom["converged"] and infom["head_residual"] <= HEAD_TOL

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    errs = dict(coef=rel(eam["coef"], ea["coef"]), q=rel(eam["q"], ea["q"]),
                constant=rel(fam["constant"], fa["constant"]), heads_m=float(np.max(np.abs(hm - h64))))
    print(f"cfg3 mixed vs fp64: {errs}; nit_fp32 {infom['nit_fp32']}, solves_fp32 {infom['solves_fp32']}")
    assert errs["coef"] < 1e-6 and errs["q"] < 1e-6 and errs["constant"] < 1e-6 and errs["heads_m"] < 1e-5
    head, _ = _oracle_rows(cfg3, fam, eam)
    print(f"cfg3 mixed: oracle head residual {head.max():.6e}")
    assert head.max() <= HEAD_TOL * (1 + 1e-6)
