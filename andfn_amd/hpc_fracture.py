"""Fracture-level evaluation of the solved network, mirroring ``andfn.hpc.hpc_fracture``.

``get_flow_nets`` and ``get_heads`` have the reference's signatures and return shapes
(``andfn/hpc/hpc_fracture.py:254-299`` and ``:302-343``): every fracture is sampled on the same
unit-disk pattern scaled by its radius, the complex potential (flow nets) or the head is
evaluated there on the GPU (``andfn_mi_omega`` / ``andfn_mi_heads``), and the points are mapped to
3D in the fracture's plane.  ``sunflower_spiral`` and ``map_2d_to_3d`` restate the reference's
pattern and plane map (``hpc_fracture.py:21-48``, ``hpc_geometry_functions.py:103-130``).
"""

import numpy as np

from .hpc_solve import get_heads as _heads_at
from .hpc_solve import get_omega as _omega_at


def sunflower_spiral(n_in, n_bnd):
    """``n_in`` points of a sunflower (golden-angle) spiral filling the unit disk, point k at radius
    sqrt((k + 1/2) / n_in) and angle pi (1 + sqrt 5) (k + 1/2), followed by ``n_bnd`` points on the
    unit circle at angles 2 pi j / (n_bnd - 1), j = 0 .. n_bnd - 1 (the first and last coincide)."""
    k = np.arange(n_in, dtype=np.float64) + 0.5
    r = np.sqrt(k / n_in)
    theta = np.pi * (1.0 + np.sqrt(5.0)) * k
    inner = r * np.cos(theta) + 1j * (r * np.sin(theta))
    return np.concatenate((inner, np.exp(1j * np.linspace(0.0, 2.0 * np.pi, n_bnd))))


def map_2d_to_3d(fracture, z):
    """Points z of a fracture's local plane as 3D coordinates x = Re z u + Im z v + centre."""
    z = np.asarray(z)
    return (np.real(z)[:, None] * fracture["x_vector"][None, :] + np.imag(z)[:, None] * fracture["y_vector"][None, :]
            + fracture["center"][None, :])


def _pattern(fracture_struc_array, z_unit):
    nf, n = len(fracture_struc_array), len(z_unit)
    radius = np.asarray(fracture_struc_array["radius"], dtype=np.float64)
    z = (z_unit[None, :] * radius[:, None]).astype(np.complex128)
    frac = np.repeat(np.arange(nf, dtype=np.int64), n)
    pnts = np.empty((nf, n, 3))
    for i in range(nf):
        pnts[i] = map_2d_to_3d(fracture_struc_array[i], z[i])
    return z, frac, pnts


def get_flow_nets(fracture_struc_array, n_points, n_boundary_points, element_struc_array, device=0):
    """Complex potential of every fracture on the sunflower pattern of ``n_points`` interior and
    ``n_boundary_points`` rim points scaled by the fracture radius.

    Returns ``(flow_nets, pnts_3d)`` shaped as the reference's: ``flow_nets`` is complex
    [n_fractures, n, n] with n = n_points + n_boundary_points, row k of fracture i holding
    Omega_i at point k (the reference's ``calc_flow_net`` assigns each point's scalar to a row of
    the per-fracture n x n array), and ``pnts_3d`` is [n_fractures, n, 3]."""
    z_unit = sunflower_spiral(n_points, n_boundary_points)
    n = len(z_unit)
    z, frac, pnts = _pattern(fracture_struc_array, z_unit)
    om = _omega_at(fracture_struc_array, element_struc_array, frac, z.ravel(), device=device).reshape(-1, n)
    flow_nets = np.repeat(om[:, :, None], n, axis=2)
    return flow_nets, pnts


def get_heads(fracture_struc_array, element_struc_array, z_array, device=0):
    """Heads of every fracture at ``z_array`` (unit-disk coordinates) scaled by the fracture radius.

    Returns ``(heads, pnts_3d)``: heads [n_fractures, n] (Re Omega / t) and the points in 3D."""
    z_unit = np.asarray(z_array, dtype=np.complex128)
    n = len(z_unit)
    z, frac, pnts = _pattern(fracture_struc_array, z_unit)
    h = _heads_at(fracture_struc_array, element_struc_array, frac, z.ravel(), device=device).reshape(-1, n)
    return h, pnts
