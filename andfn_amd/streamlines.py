"""Streamline tracking on a solved network, many lines at once on the GPU.

Follows ``DFN.streamline_tracking`` (``andfn/dfn.py:2638-2725``) line by line, but advances every
line together: each round traces one segment of every active line in its current fracture on the
device (``andfn_mi_state_trace``: the predictor-corrector steps of ``runge_kutta``,
``dfn.py:2779-2834``, and the exit tests of ``check_streamline_exit``, ``dfn.py:2727-2742``), then
moves the lines that left through an intersection into the neighbouring fracture
(``get_exit_intersection``, ``dfn.py:2744-2777``) with one batched evaluation of w on the device.

The discharge vector is the object model's ``Fracture.calc_w`` (``andfn/fracture.py:453-477``) and
the speed ``|w| / aperture`` (``fracture.py:479-497``), as in the reference.  Apertures are not part
of the solver records (``fracture_dtype_hpc``, ``andfn/element.py:146-159``), so they are passed
in, one per fracture.

One difference is deliberate: the reference's outer loop has no bound, so a line that kept
crossing intersections at its first step would never end.  Here a line stops after
``max(max_length, 1) + 1`` segments; every real segment holds at least two points, so the bound
only cuts such a line once it holds more than twice ``max_length`` points.
"""

import numpy as np

from .dtypes import INTERSECTION


def map_z_line_to_chi(z, endpoints):
    """Exterior map of a line onto the unit circle (``andfn/hpc/hpc_geometry_functions.py:11-31``)."""
    e0, e1 = endpoints[..., 0], endpoints[..., 1]
    big_z = (2 * z - e0 - e1) / (e1 - e0)
    return big_z + np.sqrt(big_z - 1) * np.sqrt(big_z + 1)


def map_chi_to_z_line(chi, endpoints):
    """Inverse of :func:`map_z_line_to_chi` (``andfn/geometry_functions.py:57-82``)."""
    e0, e1 = endpoints[..., 0], endpoints[..., 1]
    big_z = 0.5 * (chi + 1 / chi)
    return 0.5 * (big_z * (e1 - e0) + e0 + e1)


def _to_3d(fa, f, z):
    return (np.real(z)[:, None] * fa["x_vector"][f] + np.imag(z)[:, None] * fa["y_vector"][f] + fa["center"][f])


def _to_2d(fa, f, p):
    d = p - fa["center"][f]
    return np.einsum("ij,ij->i", d, fa["x_vector"][f]) + 1j * np.einsum("ij,ij->i", d, fa["y_vector"][f])


def exit_intersection(net, fa, ea, z3, elem, frac_new, frac_old, elevation, dchi=1e-4):
    """Where lines that reached intersection ``elem`` at ``z3`` (in ``frac_old``) continue in
    ``frac_new`` (``get_exit_intersection``, ``dfn.py:2744-2777``), for arrays of lines.

    The two candidate points just off the intersection on either side in the new fracture split
    the line's elevation in the ratio of their |w|; the side is chosen by the old fracture's normal.
    Returns (z0, z_start, elevation) of the next segments."""
    z3d = _to_3d(fa, frac_old, z3)
    first = frac_new == ea["frac0"][elem]
    ep = np.where(first[:, None], ea["endpoints0"][elem], ea["endpoints1"][elem])
    z = _to_2d(fa, frac_new, z3d)
    chi0 = map_z_line_to_chi(z, ep)
    chi1 = np.conj(chi0)
    za = map_chi_to_z_line(chi0 * (1 + dchi), ep)
    zb = map_chi_to_z_line(chi1 * (1 + dchi), ep)
    n = len(z3)
    w = net.state_w(np.concatenate([frac_new, frac_new]), np.concatenate([za, zb]), form="object")
    abs_a, abs_b = np.abs(w[:n]), np.abs(w[n:])
    with np.errstate(invalid="ignore", divide="ignore"):
        divide = abs_a / (abs_a + abs_b)
        nz0 = np.einsum("ij,ij->i", _to_3d(fa, frac_new, za) - z3d, fa["normal"][frac_old])
        below = nz0 < 0  # then up = zb, down = za
        up = np.where(below, zb, za)
        down = np.where(below, za, zb)
        divide = np.where(below, divide, 1 - divide)
        low = elevation < divide
        new_elev = np.where(low, elevation / divide, (elevation - divide) / (1 - divide))
    return np.where(low, down, up), za, new_elev


def track(net, fracture_struc_array, element_struc_array, z0, frac, elevation, ds, max_length, backward, aperture):
    """Trace lines i = 0..n-1 from z0[i] on fracture frac[i] with elevation[i].

    ``net`` is a :class:`andfn_amd.hpc_solve.DeviceNetwork` holding the solved state of the records.
    Returns, per line, (segments, segment fractures, speeds, exit element): segments and speeds are
    lists of arrays (one per fracture visited, as the lists ``streamline`` and ``velocity`` of the
    reference), the exit element an element index or -1 (the reference's ``False``)."""
    fa, ea = fracture_struc_array, element_struc_array
    z0 = np.atleast_1d(np.asarray(z0, dtype=np.complex128)).copy()
    n = len(z0)
    cur_f = np.broadcast_to(np.asarray(frac, dtype=np.int64), (n,)).copy()
    elev = np.broadcast_to(np.asarray(elevation, dtype=np.float64), (n,)).copy()
    ap = np.ascontiguousarray(aperture, dtype=np.float64)
    if ap.shape != (len(fa),):
        raise ValueError("aperture needs one value per fracture")
    if np.any(cur_f < 0) or np.any(cur_f >= len(fa)):
        raise ValueError("fracture index out of range")
    zs = z0.copy()
    length = np.zeros(n, dtype=np.int64)
    element = np.full(n, -1, dtype=np.int64)
    nseg = np.zeros(n, dtype=np.int64)
    segs = [[] for _ in range(n)]
    segf = [[] for _ in range(n)]
    segv = [[] for _ in range(n)]
    active = np.ones(n, dtype=bool)
    seg_cap = max(int(max_length), 1) + 1
    while np.any(active):
        idx = np.nonzero(active)[0]
        pts, vel, npts, ex = net.trace_segments(cur_f[idx], z0[idx], zs[idx], int(max_length) - length[idx], ap, ds,
                                                backward)
        ok = npts >= 0
        for k in np.nonzero(ok)[0]:
            i = idx[k]
            m = int(npts[k])
            segs[i].append(pts[k, :m].copy())
            segv[i].append(vel[k, :m].copy())
            segf[i].append(int(cur_f[i]))
        li = idx[ok]
        length[li] += npts[ok]
        element[li] = ex[ok]
        nseg[li] += 1
        active[idx[~ok]] = False  # NaN at the first step: the reference leaves its loop (dfn.py:2682-2683)
        cont = ok & (ex >= 0)
        cont[cont] = ea["_type"][ex[cont]] == INTERSECTION
        active[idx[ok & ~cont]] = False
        go = idx[cont]
        stop = nseg[go] >= seg_cap
        active[go[stop]] = False
        go = go[~stop]
        if len(go) == 0:
            continue
        rows = np.nonzero(cont)[0][~stop]
        z3 = pts[rows, npts[rows] - 1]
        e = element[go]
        f_old = cur_f[go]
        f_new = np.where(ea["frac0"][e] == f_old, ea["frac1"][e], ea["frac0"][e])
        z0[go], zs[go], elev[go] = exit_intersection(net, fa, ea, z3, e, f_new, f_old, elev[go])
        cur_f[go] = f_new
    return segs, segf, segv, element


def get_travel_time_and_length(streamline, velocity):
    """Travel time sum |dz| / v and length sum |dz| of a streamline (``dfn.py:2836-2883``):
    a list of per-fracture segments, or one segment."""
    if len(streamline) and isinstance(streamline[0], (list, np.ndarray)) and np.ndim(streamline[0]) > 0:
        time = 0.0
        length = 0.0
        for s, v in zip(streamline, velocity):
            t, le = _length_time(np.asarray(s), np.asarray(v))
            time += t
            length += le
        return time, length
    return _length_time(np.asarray(streamline), np.asarray(velocity))


def _length_time(s, v):
    # NaN points dropped first (dfn.py:2877-2883, 2886-2893)
    mask = ~np.isnan(s) & ~np.isnan(v)
    s = s[mask]
    v = v[mask]
    d = np.abs(s[1:] - s[:-1])
    return np.sum(d / v[:-1]), np.sum(d)
