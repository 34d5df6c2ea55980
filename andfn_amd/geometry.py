"""Geometry of the DFN inputs: fracture planes, disk-disk intersections and boundary lines.

These functions produce the solver's inputs (element endpoints in each fracture's local
complex plane).  They follow AnDFN's constructors so that a DFN built here yields the same
element records as AnDFN:

* fracture basis vectors: ``andfn/fracture.py:61-67``
* 3-D <-> fracture-plane maps: ``andfn/geometry_functions.py:161-220``
* plane/disk intersection with its candidate ordering: ``andfn/geometry_functions.py:302-356``
* line/circle intersection (root order (-b + sqrt) first): ``andfn/geometry_functions.py:359-402``
* element shortening: ``andfn/geometry_functions.py:974-993``

All functions are vectorised over a leading batch axis so that the pair search of a
20,000-fracture network runs in numpy rather than in a Python loop.
"""

import numpy as np


def fracture_basis(normal):
    """Unit normal and in-plane basis (x_vector, y_vector) for raw normals of shape (..., 3).

    x = normalise(n_raw x (n_raw + e_x)), falling back to n_raw + (1, 1, 1) when that cross
    product vanishes; y = normalise(n_raw x x); the stored normal is n_raw / |n_raw|
    (``andfn/fracture.py:61-67``).
    """
    n_raw = np.asarray(normal, dtype=np.float64)
    nrm = n_raw / np.linalg.norm(n_raw, axis=-1, keepdims=True)
    x = np.cross(n_raw, n_raw + np.array([1.0, 0.0, 0.0]))
    xn = np.linalg.norm(x, axis=-1, keepdims=True)
    alt = np.cross(n_raw, n_raw + np.array([1.0, 1.0, 1.0]))
    x = np.where(xn == 0, alt, x)
    x = x / np.linalg.norm(x, axis=-1, keepdims=True)
    y = np.cross(n_raw, x)
    y = y / np.linalg.norm(y, axis=-1, keepdims=True)
    return nrm, x, y


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def map_3d_to_2d(point, center, xv, yv):
    d = point - center
    return _dot(d, xv) + 1j * _dot(d, yv)


def map_2d_to_3d(z, center, xv, yv):
    z = np.asarray(z)
    return np.real(z)[..., None] * xv + np.imag(z)[..., None] * yv + center


def line_circle_intersection(z0, z1, radius):
    """Intersections of the line through z0, z1 with the circle |z| = radius (batched).

    Returns (a, b, ok).  ``ok`` is False where the discriminant is negative.  For a
    vertical line the square root of a negative number gives NaN points with ok True,
    exactly as the scalar original does.
    """
    z0 = np.asarray(z0, dtype=np.complex128)
    z1 = np.asarray(z1, dtype=np.complex128)
    radius = np.asarray(radius, dtype=np.float64)
    dx = np.real(z1 - z0)
    dy = np.imag(z1 - z0)
    vertical = dx == 0
    with np.errstate(divide="ignore", invalid="ignore"):
        m = dy / np.where(vertical, 1.0, dx)
        x0 = np.imag(z0) - m * np.real(z0)
        a = 1 + m**2
        b = 2 * x0 * m
        c = x0**2 - radius**2
        disc = b**2 - 4 * a * c
        sq = np.sqrt(np.where(disc < 0, 0.0, disc))
        x1 = (-b + sq) / (2 * a)
        x2 = (-b - sq) / (2 * a)
        y1 = m * x1 + x0
        y2 = m * x2 + x0
        xv = np.real(z0)
        yv = np.sqrt(radius**2 - xv**2)
    za = np.where(vertical, xv + 1j * yv, x1 + 1j * y1)
    zb = np.where(vertical, xv - 1j * yv, x2 + 1j * y2)
    ok = vertical | (disc >= 0)
    return za, zb, ok


def fracture_intersection(c0, n0, x0v, y0v, r0, c1, n1, x1v, y1v, r1):
    """Chord shared by two disks, in both disks' local planes (batched over pairs).

    Arguments are per-pair arrays (centre, unit normal, x/y basis, radius) of disk 0 and
    disk 1.  Returns (endpoints0, endpoints1, ok) with endpoints of shape (P, 2) complex.
    The two endpoints are the first two candidates, in the order [a0, b0, a1, b1], that lie
    in both closed disks (tolerance 1e-8); pairs with another number of valid candidates, or
    parallel planes (|n0 x n1|^2 < 1e-14), give ok False.
    """
    d = np.cross(n0, n1)
    d2 = _dot(d, d)
    par = d2 < 1e-14
    d2s = np.where(par, 1.0, d2)
    cc0 = _dot(n0, c0)
    cc1 = _dot(n1, c1)
    p0 = (cc0[:, None] * np.cross(n1, d) + cc1[:, None] * np.cross(d, n0)) / d2s[:, None]
    p1 = p0 + d
    z00 = map_3d_to_2d(p0, c0, x0v, y0v)
    z01 = map_3d_to_2d(p1, c0, x0v, y0v)
    z10 = map_3d_to_2d(p0, c1, x1v, y1v)
    z11 = map_3d_to_2d(p1, c1, x1v, y1v)
    a0, b0, ok0 = line_circle_intersection(z00, z01, r0)
    a1, b1, ok1 = line_circle_intersection(z10, z11, r1)
    cand = np.stack(
        [
            map_2d_to_3d(a0, c0, x0v, y0v),
            map_2d_to_3d(b0, c0, x0v, y0v),
            map_2d_to_3d(a1, c1, x1v, y1v),
            map_2d_to_3d(b1, c1, x1v, y1v),
        ],
        axis=1,
    )  # (P, 4, 3)
    dist0 = np.linalg.norm(cand - c0[:, None, :], axis=-1)
    dist1 = np.linalg.norm(cand - c1[:, None, :], axis=-1)
    valid = (dist0 <= r0[:, None] + 1e-8) & (dist1 <= r1[:, None] + 1e-8)
    nvalid = valid.sum(axis=1)
    ok = (~par) & ok0 & ok1 & (nvalid == 2)
    # first and second valid candidate
    order = np.argsort(~valid, axis=1, kind="stable")
    i0 = order[:, 0]
    i1 = order[:, 1]
    rows = np.arange(len(i0))
    q0 = cand[rows, i0]
    q1 = cand[rows, i1]
    e0 = np.stack([map_3d_to_2d(q0, c0, x0v, y0v), map_3d_to_2d(q1, c0, x0v, y0v)], axis=1)
    e1 = np.stack([map_3d_to_2d(q0, c1, x1v, y1v), map_3d_to_2d(q1, c1, x1v, y1v)], axis=1)
    return e0, e1, ok


def shorten_line(endpoints, se_factor):
    """Shrink segments (..., 2) about their midpoints by ``se_factor``."""
    endpoints = np.asarray(endpoints, dtype=np.complex128)
    z0 = endpoints[..., 0]
    z1 = endpoints[..., 1]
    center = (z0 + z1) / 2
    return np.stack([center + (z0 - center) * se_factor, center + (z1 - center) * se_factor], axis=-1)


def clip_to_disks(e0, r0, e1=None, r1=None, rho=1.0):
    """Clip segments to the part that lies within rho * radius of their disk centre(s).

    e0 (P, 2) are segments in disk 0's plane (radius r0); e1/r1 optionally give the same
    segments in a second disk's plane (an intersection seen from both fractures): the same
    fraction of the segment is kept in both planes.  Returns (e0', e1', ok); ok is False where
    nothing of a segment stays inside.
    """
    def interval(e, r):
        a = e[:, 0]
        d = e[:, 1] - e[:, 0]
        A = np.abs(d) ** 2
        B = np.real(a * np.conj(d))
        C = np.abs(a) ** 2 - (rho * r) ** 2
        disc = B * B - A * C
        sq = np.sqrt(np.maximum(disc, 0.0))
        As = np.where(A > 0, A, 1.0)
        lo = (-B - sq) / As
        hi = (-B + sq) / As
        return lo, hi, (disc >= 0) & (A > 0)

    lo, hi, ok = interval(e0, r0)
    if e1 is not None:
        lo1, hi1, ok1 = interval(e1, r1)
        lo, hi, ok = np.maximum(lo, lo1), np.minimum(hi, hi1), ok & ok1
    lo = np.clip(lo, 0.0, 1.0)
    hi = np.clip(hi, 0.0, 1.0)
    ok = ok & (hi - lo > 1e-9)

    def cut(e):
        d = e[:, 1] - e[:, 0]
        return np.stack([e[:, 0] + lo * d, e[:, 0] + hi * d], axis=1)

    return cut(e0), (cut(e1) if e1 is not None else None), ok


def candidate_pairs(centers, radii):
    """All pairs (i < k) whose bounding spheres overlap, |c_i - c_k| <= r_i + r_k, sorted.

    Fractures are grouped into radius classes (doubling widths) and each class pair is
    searched with a KD-tree at the class-pair's largest possible reach, then filtered
    exactly.  The result is the same set an all-pairs loop would give.
    """
    from scipy.spatial import cKDTree

    n = len(radii)
    if n < 2:
        return np.zeros((0, 2), dtype=np.int64)
    rmin = float(radii.min())
    cls = np.floor(np.log2(radii / rmin)).astype(np.int64)
    out = []
    classes = np.unique(cls)
    groups = {c: np.nonzero(cls == c)[0] for c in classes}
    trees = {c: cKDTree(centers[groups[c]]) for c in classes}
    rmax = {c: float(radii[groups[c]].max()) for c in classes}
    for ia, a in enumerate(classes):
        for b in classes[ia:]:
            ga, gb = groups[a], groups[b]
            reach = rmax[a] + rmax[b]
            pairs = trees[a].query_ball_tree(trees[b], reach)
            ii = np.repeat(np.arange(len(ga)), [len(p) for p in pairs])
            if len(ii) == 0:
                continue
            kk = np.concatenate([np.asarray(p, dtype=np.int64) for p in pairs if len(p)])
            gi = ga[ii]
            gk = gb[kk]
            keep = gi != gk
            gi, gk = gi[keep], gk[keep]
            dist = np.linalg.norm(centers[gi] - centers[gk], axis=1)
            keep = dist <= radii[gi] + radii[gk]
            gi, gk = gi[keep], gk[keep]
            lo = np.minimum(gi, gk)
            hi = np.maximum(gi, gk)
            out.append(np.stack([lo, hi], axis=1))
    if not out:
        return np.zeros((0, 2), dtype=np.int64)
    allp = np.unique(np.concatenate(out, axis=0), axis=0)
    return allp
