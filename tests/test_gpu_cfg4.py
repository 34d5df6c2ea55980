"""Adjoint at the specified size: the 5,000-fracture network (4,656 fractures kept).

Objective J = Q(x = 0) (signed total discharge of the constant-head lines on x = 0) + the sum of
the heads at 100 random observation points on fractures; gradient with respect to all 4,656
log10 transmissivities and the two boundary heads by the implicit FP64 adjoint (GMRES).

The forward solves run a fixed 400 iterations at the fixed coefficient counts (the change-based
stop rule off): the iteration contracts by ~0.92 per step on this network, so 400 steps take it
to rounding (head residual at the floor of ~1e-12 m), and finite differences are not polluted by
where a tolerance stops.  Checks:

* central differences, step 1e-4 in log10 T, on 20 fractures (the 5 largest gradients and 15
  drawn at random): within 1e-5 of max |gradient| (O(h^2) truncation ~1e-8);
* the boundary heads: J is linear in them, so the gradient equals the difference quotient of a
  0.25 m step to 1e-8, and the two head gradients sum to the 100 observation points (raising
  every prescribed head by d raises every head by d and leaves the discharges unchanged) to 1e-9;
* the adjoint costs at most twice one forward solve at head residual 1e-6 (reported ratio).
"""

import time

import numpy as np
import pytest

from andfn_amd.autograd import DFNProblem, boundary_groups
from andfn_amd.generator import build_network, workload

pytestmark = pytest.mark.gpu
NOBS = 100


@pytest.fixture(scope="module")
def problem():
    print("cfg4: building the 5,000-fracture network", flush=True)
    d = build_network(workload("cfg2", constants=dict(MAX_ERROR=0.0, MAX_ITERATIONS=400)))
    d.consolidate_dfn()
    groups = boundary_groups(d)
    fa = d.fractures_struc_array_hpc
    rng = np.random.default_rng(41)
    of = rng.integers(0, len(fa), NOBS)
    oz = fa["radius"][of] * rng.uniform(0.1, 0.7, NOBS) * np.exp(1j * rng.uniform(0, 2 * np.pi, NOBS))
    p = DFNProblem(d, groups, of, oz, head_tol=0.0, adjoint_tol=1e-12, fixed_iterations=True)
    lt = np.log10(p.t0)
    gh = np.array([1.0, 0.0])
    Q, h = p.evaluate(lt, gh)
    print(f"cfg4: forward head residual {p.last_info['head_residual']:.3e}", flush=True)
    t0 = time.perf_counter()
    g_lt, g_h = p.gradients(lt, gh, np.array([1.0, 0.0]), np.ones(NOBS))
    t_adj = time.perf_counter() - t0
    print(f"cfg4: adjoint {t_adj:.3f}s, {p.last_adjoint_info['nit']} iterations", flush=True)
    yield dict(p=p, lt=lt, gh=gh, J=Q[0] + h.sum(), g_lt=g_lt, g_h=g_h, t_adj=t_adj)
    p.close()


def _J(p, lt, gh):
    Q, h = p.evaluate(lt, gh)
    return Q[0] + h.sum()


def test_cfg4_log10_t_gradient_matches_central_differences(problem):
    p, lt, gh, g = problem["p"], problem["lt"], problem["gh"], problem["g_lt"]
    assert p.last_info["head_residual"] < 1e-10
    s = np.max(np.abs(g))
    rng = np.random.default_rng(43)
    pick = list(np.argsort(-np.abs(g))[:5]) + list(rng.choice(len(lt), 15, replace=False))
    step = 1e-4
    worst = 0.0
    for f in pick:
        lp, lm = lt.copy(), lt.copy()
        lp[f] += step
        lm[f] -= step
        fd = (_J(p, lp, gh) - _J(p, lm, gh)) / (2 * step)
        worst = max(worst, abs(fd - g[f]) / s)
        print(f"fracture {f}: central difference {fd:.10e}, adjoint {g[f]:.10e}", flush=True)
        assert abs(fd - g[f]) <= 1e-5 * s, (f, fd, g[f])
    print(f"cfg4: worst |FD - adjoint| / max|gradient| = {worst:.2e} over {len(pick)} fractures")


def test_cfg4_boundary_head_gradients(problem):
    p, lt, gh, g_h = problem["p"], problem["lt"], problem["gh"], problem["g_h"]
    assert abs(g_h.sum() - NOBS) <= 1e-9 * NOBS
    dh = 0.25
    for k in range(2):
        e = np.zeros(2)
        e[k] = dh
        fd = (_J(p, lt, gh + e) - _J(p, lt, gh - e)) / (2 * dh)
        print(f"head {k}: difference quotient {fd:.12e}, adjoint {g_h[k]:.12e}", flush=True)
        assert abs(fd - g_h[k]) <= 1e-8 * np.max(np.abs(g_h))


def test_cfg4_adjoint_cost_against_a_forward_solve(problem):
    d = build_network(workload("cfg2"))
    d.consolidate_dfn()
    p = DFNProblem(d, boundary_groups(d), problem["p"].obs_frac, problem["p"].obs_z, head_tol=1e-6)
    try:
        lt = np.log10(p.t0)
        gh = np.array([1.0, 0.0])
        p.evaluate(lt, gh)  # warm-up
        t0 = time.perf_counter()
        p.evaluate(lt, gh)
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        p.gradients(lt, gh, np.array([1.0, 0.0]), np.ones(NOBS))
        t_adj = time.perf_counter() - t0
    finally:
        p.close()
    print(f"cfg4: forward {t_fwd:.3f}s (head residual 1e-6), adjoint {t_adj:.3f}s, ratio {t_adj / t_fwd:.2f}")
    assert t_adj <= 2.0 * t_fwd
