"""Run-ahead iterations: one HIP graph launch per iteration, the head-residual stop decided on the
device (the host keeps one iteration enqueued ahead of the residual it reads).

The loop of ``andfn.hpc.hpc_solve.solve`` (hpc_solve.py:152-239, with the head-residual stop
rule) must give the same iterates whichever way it is driven: the graph path, its launches made
directly (ANDFN_MI_NO_GRAPH) and the host-driven loop (ANDFN_MI_NO_RUNAHEAD) are compared bit for
bit, for a run stopped by the tolerance and for one that reaches MAX_ITERATIONS (with AnDFN's
extra sweep at the cap).
"""

import numpy as np
import pytest

from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import last_solve_info, solve

pytestmark = pytest.mark.gpu


def _records(name, **consts):
    d = build_network(workload(name, constants=consts))
    d.consolidate_dfn()
    return d.fractures_struc_array_hpc, d.elements_struc_array_hpc, d.discharge_int, d.constants


def _run(monkeypatch, env, recs, **kw):
    for k in ("ANDFN_MI_NO_RUNAHEAD", "ANDFN_MI_NO_GRAPH"):
        monkeypatch.delenv(k, raising=False)
    if env:
        monkeypatch.setenv(env, "1")
    fa0, ea0, K, consts = recs
    fa, ea = fa0.copy(), ea0.copy()
    solve(fa, ea, K, consts, **kw)
    return fa, ea, dict(last_solve_info())


@pytest.mark.parametrize("max_it,tol", [(3000, 1e-9), (25, 1e-14)])
def test_runahead_iterates_equal_the_host_loop(monkeypatch, max_it, tol):
    recs = _records("cfg1", MAX_ERROR=0.0, MAX_ITERATIONS=max_it)
    runs = {env: _run(monkeypatch, env, recs, precision="fp64", head_tol=tol)
            for env in (None, "ANDFN_MI_NO_GRAPH", "ANDFN_MI_NO_RUNAHEAD")}
    fg, eg, ig = runs[None]
    print({k: ig[k] for k in ("nit", "head_residual", "runahead_iterations", "graph_iterations")})
    assert ig["runahead_iterations"] == ig["nit"] - 1
    # the first use of each of the three per-iteration graphs runs its launches directly
    assert ig["graph_iterations"] >= ig["runahead_iterations"] - 3
    assert runs["ANDFN_MI_NO_GRAPH"][2]["graph_iterations"] == 0
    assert runs["ANDFN_MI_NO_RUNAHEAD"][2]["runahead_iterations"] == 0
    for env in ("ANDFN_MI_NO_GRAPH", "ANDFN_MI_NO_RUNAHEAD"):
        f, e, i = runs[env]
        assert i["nit"] == ig["nit"] and i["converged"] == ig["converged"], (env, i["nit"], ig["nit"])
        assert i["head_residual"] == ig["head_residual"]
        for name in ("coef", "q", "ncoef", "error", "error_coef"):
            assert np.array_equal(e[name], eg[name]), (env, name)
        assert np.array_equal(f["constant"], fg["constant"]), env
    if max_it == 3000:
        assert ig["converged"] and ig["nit"] < max_it and ig["head_residual"] <= tol
    else:
        assert not ig["converged"] and ig["nit"] == max_it


def test_runahead_mixed_form_converges_like_the_host_loop(monkeypatch):
    recs = _records("cfg1", MAX_ERROR=0.0, MAX_ITERATIONS=3000)
    fg, eg, ig = _run(monkeypatch, None, recs, precision="mixed", head_tol=1e-9)
    fh, eh, ih = _run(monkeypatch, "ANDFN_MI_NO_RUNAHEAD", recs, precision="mixed", head_tol=1e-9)
    print(ig["nit"], ig["nit_fp32"], ih["nit"], ih["nit_fp32"])
    assert ig["converged"] and ih["converged"] and ig["nit_fp32"] > 0
    assert ig["head_residual"] <= 1e-9 and ig["graph_iterations"] > 0
    # the phase switch follows the residual one iteration later when run ahead: same fixed point
    s = np.max(np.abs(eh["coef"]))
    assert np.max(np.abs(eg["coef"] - eh["coef"])) <= 1e-6 * s
    assert np.max(np.abs(eg["q"] - eh["q"])) <= 1e-6 * np.max(np.abs(eh["q"]))
