"""Streamline tracking on a solved network, many lines at once on the GPU.

``track`` is ``DFN.streamline_tracking`` (``andfn/dfn.py:2638-2725``) for arrays of start points:
every line is traced on the device (``andfn_mi_state_streamlines``), one lane per line, segment by
segment: the predictor-corrector steps of ``runge_kutta`` (``dfn.py:2779-2834``), the exit tests of
``check_streamline_exit`` (``dfn.py:2727-2742``) and the move into the neighbouring fracture of
``get_exit_intersection`` (``dfn.py:2744-2777``) all run in the kernel, so a line costs no round trip
to the host per segment.

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


def track(net, fracture_struc_array, element_struc_array, z0, frac, elevation, ds, max_length, backward, aperture):
    """Trace lines i = 0..n-1 from z0[i] on fracture frac[i] with elevation[i].

    ``net`` is a :class:`andfn_amd.hpc_solve.DeviceNetwork` holding the solved state of the records.
    Returns, per line, (segments, segment fractures, speeds, exit element): segments and speeds are
    lists of arrays (one per fracture visited, as the lists ``streamline`` and ``velocity`` of the
    reference), the exit element an element index or -1 (the reference's ``False``)."""
    fa = fracture_struc_array
    z0 = np.atleast_1d(np.asarray(z0, dtype=np.complex128))
    n = len(z0)
    fr = np.broadcast_to(np.asarray(frac, dtype=np.int64), (n,))
    el = np.broadcast_to(np.asarray(elevation, dtype=np.float64), (n,))
    ap = np.ascontiguousarray(aperture, dtype=np.float64)
    if ap.shape != (len(fa),):
        raise ValueError("aperture needs one value per fracture")
    pts, vel, npts, sf, sp, nseg, ex = net.streamlines(fr, z0, el, ap, ds, max_length, backward)
    segs, segf, segv = [], [], []
    p = 0
    q = 0
    for i in range(n):
        ls, lf, lv = [], [], []
        for j in range(int(nseg[i])):
            m = int(sp[q])
            ls.append(pts[p:p + m])
            lv.append(vel[p:p + m])
            lf.append(int(sf[q]))
            p += m
            q += 1
        segs.append(ls)
        segf.append(lf)
        segv.append(lv)
    return segs, segf, segv, ex


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
