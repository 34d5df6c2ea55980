"""Implicit adjoint of the converged solve against finite differences and exact identities.

Objective: J = Q_x0 (total discharge of the constant-head lines on x = 0) + sum of the heads at
20 observation points, on the 200-fracture network (growth off: its MAX_NCOEF equals its
coefficient count, so J is a smooth function of the parameters).  The forward solves run a
fixed 3000 iterations (the change-based stop rule disabled), which converges the damped
iteration (contraction ~0.986 per iteration here) to rounding, so finite differences are not
polluted by where a tolerance happens to stop.

Checks and tolerances:
  * GMRES and the transposed damped iteration agree (1e-8 relative to max |gradient|);
  * J is linear in the prescribed heads: the gradient equals the exact difference quotient
    (1e-8 relative), and raising every prescribed head by d raises every head by d and leaves the
    discharges unchanged, so the head gradients sum to the number of observation points (1e-9);
  * d J / d log10 T against central differences with step 1e-4 (1e-5 relative to max |gradient|;
    the O(h^2) truncation is ~1e-8);
  * the torch.autograd.Function returns the adjoint's gradients.
"""

import math

import numpy as np
import pytest
import torch

from andfn_amd import _lib
from andfn_amd.autograd import DFNProblem, DFNSolve, boundary_groups
from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork

pytestmark = pytest.mark.gpu
NOBS = 20


def _build(name, **consts):
    d = build_network(workload(name, constants=consts))
    d.consolidate_dfn()
    return d


def _obs(fa, n, seed):
    rng = np.random.default_rng(seed)
    of = rng.integers(0, len(fa), n)
    oz = fa["radius"][of] * rng.uniform(0.1, 0.7, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return of, oz


@pytest.fixture(scope="module")
def problem():
    print("adjoint fixture cfg1, 3000 iterations", flush=True)
    d = _build("cfg1", MAX_ERROR=0.0, MAX_ITERATIONS=3000)
    groups = boundary_groups(d)
    of, oz = _obs(d.fractures_struc_array_hpc, NOBS, 5)
    p = DFNProblem(d, groups, of, oz, head_tol=0.0, adjoint_tol=1e-11, fixed_iterations=True)
    lt = np.log10(p.t0)
    gh = np.array([1.0, 0.0])
    Q, h = p.evaluate(lt, gh)
    dq = np.zeros(len(p.ea))
    dq[p.groups[0]] = 1.0
    gt, ghd, info = p.net.adjoint(dq=dq, obs_frac=of, obs_z=oz, dhead=np.ones(NOBS), tol=1e-11,
                                  accept_unconverged=True)
    assert info["converged"], info
    yield dict(p=p, lt=lt, gh=gh, Q=Q, h=h, dq=dq, gt=gt, ghd=ghd, info=info, of=of, oz=oz)
    p.close()


def _J(p, lt, gh):
    Q, h = p.evaluate(lt, gh)
    return Q[0] + h.sum()


def test_objective_value_matches_forward(problem):
    J = problem["Q"][0] + problem["h"].sum()
    assert abs(problem["info"]["objective"] - J) <= 1e-12 * abs(J)


def test_gmres_matches_transposed_fixed_point(problem):
    p = problem["p"]
    p.evaluate(problem["lt"], problem["gh"])
    gt2, gh2, info = p.net.adjoint(dq=problem["dq"], obs_frac=problem["of"], obs_z=problem["oz"],
                                   dhead=np.ones(NOBS), tol=1e-12, method="fixed_point", accept_unconverged=True)
    assert info["converged"]
    s = np.max(np.abs(problem["gt"]))
    assert np.max(np.abs(gt2 - problem["gt"])) <= 1e-8 * s
    assert np.max(np.abs(gh2 - problem["ghd"])) <= 1e-8 * np.max(np.abs(problem["ghd"]))


def test_head_gradients_sum_to_observation_count(problem):
    p, ghd = problem["p"], problem["ghd"]
    total = sum(ghd[idx].sum() for idx in p.groups)
    assert abs(total - NOBS) <= 1e-9 * NOBS
    # only constant-head elements carry a head gradient
    ty = p.ea["_type"]
    assert np.all(ghd[(ty != 2) & (ty != 3)] == 0.0)


def test_head_gradient_is_the_exact_difference_quotient(problem):
    p, lt = problem["p"], problem["lt"]
    dh = 0.25
    Jp = _J(p, lt, problem["gh"] + np.array([dh, 0.0]))
    Jm = _J(p, lt, problem["gh"] - np.array([dh, 0.0]))
    fd = (Jp - Jm) / (2 * dh)
    adj = problem["ghd"][p.groups[0]].sum()
    assert abs(fd - adj) <= 1e-8 * abs(adj)


def test_log10_t_gradient_matches_central_differences(problem):
    p, lt = problem["p"], problem["lt"]
    adj = problem["gt"] * p.t0 * math.log(10.0)
    s = np.max(np.abs(adj))
    rng = np.random.default_rng(11)
    pick = list(np.argsort(-np.abs(adj))[:2]) + list(rng.integers(0, len(lt), 1))
    h = 1e-4
    for f in pick:
        print(f"central difference, fracture {f}", flush=True)
        lp = lt.copy()
        lp[f] += h
        lm = lt.copy()
        lm[f] -= h
        fd = (_J(p, lp, problem["gh"]) - _J(p, lm, problem["gh"])) / (2 * h)
        assert abs(fd - adj[f]) <= 1e-5 * s, (f, fd, adj[f])


def test_torch_function_backward_is_the_adjoint(problem):
    p = problem["p"]
    lt = torch.tensor(problem["lt"], dtype=torch.float64, requires_grad=True)
    gh = torch.tensor(problem["gh"], dtype=torch.float64, requires_grad=True)
    Q, h = DFNSolve.apply(lt, gh, p)
    assert abs(float(Q.detach()[0]) - problem["Q"][0]) <= 1e-12 * abs(problem["Q"][0])
    loss = Q[0] + h.sum()
    loss.backward()
    ref = problem["gt"] * p.t0 * math.log(10.0)
    assert np.max(np.abs(lt.grad.numpy() - ref)) <= 1e-9 * np.max(np.abs(ref))
    ref_h = np.array([problem["ghd"][idx].sum() for idx in p.groups])
    assert np.max(np.abs(gh.grad.numpy() - ref_h)) <= 1e-9 * np.max(np.abs(ref_h))


def test_adjoint_with_wells_and_growth():
    print("adjoint dfn40w (wells, coefficient growth)", flush=True)
    d = _build("dfn40w", MAX_ITERATIONS=500)
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    of, oz = _obs(fa, 10, 9)
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        net.run(head_tol=1e-6)
        ty = ea["_type"]
        dq = (ty == 2).astype(np.float64)  # total well discharge
        g1, h1, i1 = net.adjoint(dq=dq, obs_frac=of, obs_z=oz, dhead=np.ones(10), tol=1e-12)
        g2, h2, i2 = net.adjoint(dq=dq, obs_frac=of, obs_z=oz, dhead=np.ones(10), tol=1e-12, method="fixed_point")
        assert i1["converged"] and i2["converged"]
        assert np.max(np.abs(g1 - g2)) <= 1e-8 * np.max(np.abs(g1))
        fixed = (ty == 2) | (ty == 3)
        assert np.any(h1[ty == 2] != 0.0)
        assert abs(h1[fixed].sum() - 10.0) <= 1e-8 * 10.0
    finally:
        net.close()


def test_adjoint_needs_a_solved_state():
    d = _build("dfn40")
    net = DeviceNetwork(d.fractures_struc_array_hpc, d.elements_struc_array_hpc, d.discharge_int, d.constants)
    try:
        with pytest.raises(_lib.AndfnMiError):
            net.adjoint(dq=np.ones(len(d.elements_struc_array_hpc)))
    finally:
        net.close()


def test_unconverged_forward_is_refused():
    """A forward run that stops at MAX_ITERATIONS above its head tolerance is not the fixed point:
    the C ABI refuses its adjoint (unless the caller accepts that state) and the torch problem
    raises."""
    print("adjoint refusal dfn40, 3 iterations", flush=True)
    d = _build("dfn40", MAX_ITERATIONS=3)
    fa, ea = d.fractures_struc_array_hpc, d.elements_struc_array_hpc
    net = DeviceNetwork(fa, ea, d.discharge_int, d.constants)
    try:
        info = net.run(head_tol=1e-9)
        assert not info["converged"]
        with pytest.raises(_lib.AndfnMiError, match="fixed point"):
            net.adjoint(dq=np.ones(len(ea)))
        g, h, ai = net.adjoint(dq=np.ones(len(ea)), accept_unconverged=True)
        assert np.all(np.isfinite(g))
    finally:
        net.close()
    p = DFNProblem(d, boundary_groups(d), head_tol=1e-9)
    try:
        with pytest.raises(RuntimeError, match="fixed point"):
            p.evaluate(np.log10(p.t0), np.array([1.0, 0.0]))
    finally:
        p.close()
