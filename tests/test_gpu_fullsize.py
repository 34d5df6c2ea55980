"""The cfg2 workload at full size (5,000 disks generated, 4,656 kept, 19,661 unknowns), checked
through properties that do not depend on its size.

The oracle's full solve of this network takes over a minute and a half (SuperLU alone ~6 s, 160
sweeps of ~0.5 s), so the GPU's converged state is checked instead:

* the head residual, recomputed by the oracle from the GPU's coefficients, discharges and
  fracture constants (its own discharge matrix and right-hand-side sweep), is within the stop
  tolerance and agrees with the residual the GPU reported;
* the continuity rows of the discharge system (every fracture's discharges sum to zero) hold to
  rounding;
* two FP64 runs give bit-identical results (every reduction has a fixed order);
* the mixed form stops at the same head residual within the errors stated for it;
* values: 160 iterations at the fixed coefficient counts (AnDFN's change-based stop off), the
  iteration count the head-residual stop takes on this network, give coefficients, discharges
  and fracture constants within 1e-9 of the oracle's 160 iterations (relative to the largest
  magnitude of each quantity), with identical counts.  The oracle here refines every SuperLU
  solve once: on this matrix SciPy's SuperLU solve alone is accurate to ~1.2e-10 relative (its
  residual is ~1e-10 of |b|; one refinement step, or the row-equilibrated dense LU, gives 2e-16 and
  3e-14), and over 160 iterations that solve error builds up to ~2e-9 in the state.  The test
  checks that too: the GPU's distance from the unrefined oracle is no larger than the unrefined
  oracle's own distance from the refined one (plus 1e-10).
"""

import numpy as np
import pytest

from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork, last_solve_info, solve
from oracle import oracle

pytestmark = pytest.mark.gpu

HEAD_TOL = 1e-6


@pytest.fixture(scope="module")
def cfg2():
    print("cfg2: building the network", flush=True)
    d = build_network(workload("cfg2"))
    d.consolidate_dfn()
    return d


def _run(d, precision, factor="sparse"):
    fa = d.fractures_struc_array_hpc.copy()
    ea = d.elements_struc_array_hpc.copy()
    print(f"cfg2: {precision} solve ({factor} LU)", flush=True)
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    info = net.run(precision=precision, head_tol=HEAD_TOL, factor=factor)
    net.download()
    net.close()
    return fa, ea, info


@pytest.fixture(scope="module")
def fp64(cfg2):
    return _run(cfg2, "fp64")


def _oracle_rows(d, fa, ea):
    """A y - b of the oracle's discharge system at the state held by the records, with the head
    rows scaled to metres (hpc_solve.py:832-887 for A, 155-174 for b)."""
    O = oracle.OracleHandle(fa.copy(), ea.copy(), d.discharge_int)
    y = np.zeros(O.n)
    O.lib.oracle_old_discharges(O.h, oracle._ptr(y))
    A = O.discharge_matrix()
    b = O.head_rhs()
    nde = O.nde
    O.close()
    r = A @ y - b
    de = np.nonzero(np.isin(ea["_type"], (0, 2, 3)))[0]
    scale = np.ones(nde)
    chl = ea["_type"][de] != 0
    scale[chl] = 1.0 / fa["t"][ea["frac0"][de[chl]]]
    return np.abs(r[:nde]) * scale, r[nde:], y


def test_head_residual_recomputed_by_the_oracle(cfg2, fp64):
    fa, ea, info = fp64
    assert info["converged"] and info["head_residual"] <= HEAD_TOL
    head, cont, y = _oracle_rows(cfg2, fa, ea)
    print(f"cfg2 fp64: GPU head residual {info['head_residual']:.6e}, oracle {head.max():.6e}")
    assert head.max() <= HEAD_TOL * (1 + 1e-6)
    assert abs(head.max() - info["head_residual"]) <= 1e-3 * HEAD_TOL


def test_fracture_discharges_balance(cfg2, fp64):
    fa, ea, _ = fp64
    _, cont, y = _oracle_rows(cfg2, fa, ea)
    qmax = np.max(np.abs(ea["q"]))
    print(f"cfg2 fp64: continuity rows max {np.max(np.abs(cont)):.3e} (max |q| {qmax:.3e})")
    assert np.max(np.abs(cont)) <= 1e-10 * qmax


def test_fp64_runs_are_bit_identical(cfg2, fp64):
    fa, ea, info = fp64
    fa2, ea2, info2 = _run(cfg2, "fp64")
    assert info2["nit"] == info["nit"]
    assert np.array_equal(ea2["ncoef"], ea["ncoef"])
    assert np.array_equal(ea2["coef"], ea["coef"])
    assert np.array_equal(ea2["q"], ea["q"])
    assert np.array_equal(fa2["constant"], fa["constant"])


def test_lu_stream_placement_is_bit_identical(cfg2, monkeypatch):
    """The dense tournament LU's panel kernels on reserved CUs (the default) or on plain streams
    compute the same arithmetic in the same order: results identical to the bit."""
    fa, ea, info = _run(cfg2, "fp64", "tournament")
    monkeypatch.setenv("ANDFN_MI_LU_RESERVE", "0")
    fa2, ea2, info2 = _run(cfg2, "fp64", "tournament")
    assert info2["nit"] == info["nit"]
    assert np.array_equal(ea2["coef"], ea["coef"])
    assert np.array_equal(ea2["q"], ea["q"])
    assert np.array_equal(fa2["constant"], fa["constant"])


def test_mixed_form_against_fp64(cfg2, fp64):
    fa, ea, info = fp64
    fam, eam, infom = _run(cfg2, "mixed")
    assert infom["converged"] and infom["head_residual"] <= HEAD_TOL
    assert infom["nit_fp32"] > 0

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    errs = dict(coef=rel(eam["coef"], ea["coef"]), q=rel(eam["q"], ea["q"]),
                constant=rel(fam["constant"], fa["constant"]))
    print(f"cfg2 mixed vs fp64: {errs}")
    assert errs["coef"] < 1e-7 and errs["q"] < 1e-7 and errs["constant"] < 1e-9
    head, _, _ = _oracle_rows(cfg2, fam, eam)
    assert head.max() <= HEAD_TOL * (1 + 1e-6)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _errs(fx, ex, fy, ey):
    return dict(coef=_rel(ex["coef"], ey["coef"]), q=_rel(ex["q"], ey["q"]),
                constant=_rel(fx["constant"], fy["constant"]))


@pytest.fixture(scope="module")
def fixed160(cfg2):
    """The GPU's and the refined oracle's states after 160 iterations at fixed counts (each oracle
    run takes about a minute and a half: the two comparisons are separate tests)."""
    consts = np.array(cfg2.constants, copy=True)
    consts["MAX_ERROR"] = 0.0
    consts["MAX_ITERATIONS"] = 160
    fa = cfg2.fractures_struc_array_hpc.copy()
    ea = cfg2.elements_struc_array_hpc.copy()
    fo, eo = fa.copy(), ea.copy()
    print("cfg2: GPU, 160 iterations", flush=True)
    solve(fa, ea, cfg2.discharge_int, consts)
    info = last_solve_info()
    print("cfg2: oracle, 160 iterations, refined SuperLU solves", flush=True)
    oinfo = {}
    oracle.solve(fo, eo, cfg2.discharge_int, consts, info=oinfo, refine=1)
    return dict(consts=consts, gpu=(fa, ea, info), ref=(fo, eo, oinfo))


def test_fixed_count_values_match_the_oracle(fixed160):
    fa, ea, info = fixed160["gpu"]
    fo, eo, oinfo = fixed160["ref"]
    assert info["nit"] == oinfo["nit"] == 160
    assert np.array_equal(ea["ncoef"], eo["ncoef"]) and np.array_equal(ea["nint"], eo["nint"])
    e_ref = _errs(fa, ea, fo, eo)
    print(f"cfg2 160 iterations, GPU vs oracle (refined solves): {e_ref}; head residual GPU "
          f"{info['head_residual']:.3e}, oracle {oinfo['head_residual']:.3e}", flush=True)
    assert e_ref["coef"] < 1e-9 and e_ref["q"] < 1e-9 and e_ref["constant"] < 1e-9, e_ref


def test_fixed_count_values_against_the_unrefined_oracle(cfg2, fixed160):
    fa, ea, _ = fixed160["gpu"]
    fo, eo, _ = fixed160["ref"]
    fp = cfg2.fractures_struc_array_hpc.copy()
    ep = cfg2.elements_struc_array_hpc.copy()
    print("cfg2: oracle, 160 iterations, SuperLU solves as in AnDFN", flush=True)
    oracle.solve(fp, ep, cfg2.discharge_int, fixed160["consts"])
    e_gpu = _errs(fa, ea, fp, ep)
    e_slu = _errs(fo, eo, fp, ep)
    print(f"cfg2: GPU vs unrefined oracle {e_gpu}; refined vs unrefined oracle {e_slu}")
    for k in e_gpu:
        assert e_gpu[k] <= e_slu[k] + 1e-10, (k, e_gpu, e_slu)
