"""Gradients through the DFN solve as a ``torch.autograd.Function``.

The forward pass solves the network on the GPU for per-fracture log10 transmissivities and the
prescribed heads of groups of constant-head elements (for example the two boundary planes), and
returns the total discharge of each group and the heads at observation points.  The backward
pass is the implicit adjoint of the converged fixed point (``andfn_mi_adjoint``): one transposed
solve per backward call gives the gradient of ``sum(grad_Q * Q) + sum(grad_h * h)`` with respect
to every log10 transmissivity and every group head, with the coefficient counts frozen at their
converged values.  Nothing is unrolled and no iterate is stored.

PyTorch only carries the tensors; the solve and its adjoint run in ``libandfn_mi.so``.

    problem = DFNProblem(dfn, groups={"x0": idx0, "xL": idxL}, obs_frac=f, obs_z=z)
    Q, h = DFNSolve.apply(log10_t, group_heads, problem)
    (Q[0] + h.sum()).backward()
"""

import math

import numpy as np
import torch

from .hpc_solve import DeviceNetwork

LN10 = math.log(10.0)


def boundary_groups(dfn, labels=("x0", "xL")):
    """Element indices of the constant-head lines created with each label (in record order)."""
    if dfn.elements is None:
        dfn.get_elements()
    out = {}
    for lab in labels:
        out[lab] = np.array([i for i, e in enumerate(dfn.elements)
                             if e._type == 3 and str(e.label).startswith(lab + "_")], dtype=np.int64)
    return out


class DFNProblem:
    """A network resident on the GPU together with the observables of a gradient study.

    groups: ordered mapping name -> element indices of constant-head elements sharing one head
    obs_frac, obs_z: observation points (fracture index, local complex coordinate)
    head_tol: the forward solves stop at this head residual (metres); a solve that reaches
        MAX_ITERATIONS first raises, since the implicit adjoint is the gradient of the fixed point
    fixed_iterations: instead run every forward solve for exactly MAX_ITERATIONS iterations (the
        stop rules off; for studies that converge the iteration to rounding, such as finite-
        difference checks) and accept that state for the adjoint
    """

    def __init__(self, dfn, groups, obs_frac=None, obs_z=None, precision="fp64", head_tol=1e-6, adjoint_tol=1e-10,
                 device=0, fixed_iterations=False):
        if dfn.fractures_struc_array_hpc is None:
            dfn.consolidate_dfn()
        self.fa = dfn.fractures_struc_array_hpc
        self.ea = dfn.elements_struc_array_hpc
        self.net = DeviceNetwork(self.fa, self.ea, dfn.discharge_int, dfn.constants, device=device)
        self.group_names = list(groups)
        self.groups = [np.asarray(groups[k], dtype=np.int64) for k in self.group_names]
        self.obs_frac = np.zeros(0, np.int64) if obs_frac is None else np.asarray(obs_frac, dtype=np.int64)
        self.obs_z = np.zeros(0, np.complex128) if obs_z is None else np.asarray(obs_z, dtype=np.complex128)
        self.precision = precision
        self.head_tol = head_tol
        self.adjoint_tol = adjoint_tol
        self.fixed_iterations = bool(fixed_iterations)
        if self.fixed_iterations and float(np.asarray(dfn.constants)["MAX_ERROR"].reshape(-1)[0]) != 0.0:
            raise ValueError("fixed_iterations needs MAX_ERROR = 0 (AnDFN's change-based stop rule off)")
        self.t0 = self.fa["t"].astype(np.float64).copy()
        ty = self.ea["_type"]
        fr0 = self.ea["frac0"]
        self.head0 = np.zeros(len(self.ea))
        fixed = (ty == 2) | (ty == 3)
        self.head0[fixed] = self.ea["phi"][fixed] / self.t0[fr0[fixed]]
        self._key = None
        self.last_info = None
        self.last_adjoint_info = None

    def _heads(self, group_heads):
        h = self.head0.copy()
        for g, idx in enumerate(self.groups):
            h[idx] = group_heads[g]
        return h

    def evaluate(self, log10_t, group_heads):
        """Solve for the parameters; returns (Q per group, heads at the observation points)."""
        log10_t = np.asarray(log10_t, dtype=np.float64)
        group_heads = np.asarray(group_heads, dtype=np.float64)
        self.net.set_parameters(t=10.0 ** log10_t, head=self._heads(group_heads))
        self.last_info = self.net.run(self.precision, head_tol=0.0 if self.fixed_iterations else self.head_tol)
        if not self.last_info["converged"] and not self.fixed_iterations:
            self._key = None
            raise RuntimeError(f"forward solve stopped at MAX_ITERATIONS ({self.last_info['nit']}) with head residual "
                               f"{self.last_info['head_residual']:.3e} m above head_tol {self.head_tol:g}: the implicit "
                               f"adjoint needs the fixed point")
        q, _ = self.net.state_discharges()
        Q = np.array([q[idx].sum() for idx in self.groups])
        h = self.net.state_heads(self.obs_frac, self.obs_z) if len(self.obs_frac) else np.zeros(0)
        self._key = (log10_t.tobytes(), group_heads.tobytes())
        return Q, h

    def gradients(self, log10_t, group_heads, gQ, gH):
        """d(gQ.Q + gH.h)/d(log10_t, group_heads) at the parameters (re-solves if the resident
        state belongs to other parameters)."""
        log10_t = np.asarray(log10_t, dtype=np.float64)
        group_heads = np.asarray(group_heads, dtype=np.float64)
        if self._key != (log10_t.tobytes(), group_heads.tobytes()):
            self.evaluate(log10_t, group_heads)
        dq = np.zeros(len(self.ea))
        for g, idx in enumerate(self.groups):
            dq[idx] += gQ[g]
        gt, gh, info = self.net.adjoint(dq=dq, obs_frac=self.obs_frac if len(self.obs_frac) else None,
                                        obs_z=self.obs_z if len(self.obs_frac) else None,
                                        dhead=np.asarray(gH, dtype=np.float64) if len(self.obs_frac) else None,
                                        tol=self.adjoint_tol, accept_unconverged=self.fixed_iterations)
        self.last_adjoint_info = info
        g_log10_t = gt * (10.0 ** log10_t) * LN10
        g_heads = np.array([gh[idx].sum() for idx in self.groups])
        return g_log10_t, g_heads

    def close(self):
        self.net.close()


class DFNSolve(torch.autograd.Function):
    """(log10_t [F], group_heads [G], problem) -> (Q [G], heads [n_obs]), float64."""

    @staticmethod
    def forward(ctx, log10_t, group_heads, problem):
        lt = log10_t.detach().to("cpu", torch.float64).numpy().copy()
        gh = group_heads.detach().to("cpu", torch.float64).numpy().copy()
        Q, h = problem.evaluate(lt, gh)
        ctx.problem = problem
        ctx.params = (lt, gh)
        ctx.devices = (log10_t.device, group_heads.device)
        return (torch.as_tensor(Q, dtype=torch.float64, device=log10_t.device),
                torch.as_tensor(h, dtype=torch.float64, device=log10_t.device))

    @staticmethod
    def backward(ctx, gQ, gH):
        lt, gh = ctx.params
        gq = np.zeros(len(ctx.problem.groups)) if gQ is None else gQ.detach().to("cpu", torch.float64).numpy()
        ghh = np.zeros(len(ctx.problem.obs_frac)) if gH is None else gH.detach().to("cpu", torch.float64).numpy()
        g_lt, g_heads = ctx.problem.gradients(lt, gh, gq, ghh)
        return (torch.as_tensor(g_lt, dtype=torch.float64, device=ctx.devices[0]),
                torch.as_tensor(g_heads, dtype=torch.float64, device=ctx.devices[1]), None)
