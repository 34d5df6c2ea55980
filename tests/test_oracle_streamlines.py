"""Pins of the oracle's discharge vector w and streamline tracking by what the mathematics fixes.

* The object-model form of w (``andfn/fracture.py:453-477``) is -dOmega/dz: checked against central
  differences of the real part of the oracle's Omega (itself pinned by AnDFN's golden runs), so a
  wrong sign, a missing 1/r, a wrong map derivative or a dropped q term fails.
* The jitted form (``andfn/hpc/hpc_fracture.py:139-174``) differs in three ways the reference's code
  fixes; each is checked on a network reduced to one element: the bounding circle enters as
  ``taylor_series_d1(z/r)`` = r dOmega/dz (``hpc_bounding_circle.py:97-115``), impermeable objects
  are not summed, the other elements agree with the object model.
* The series derivatives (``hpc_math_functions.py:75-100, 163-190``) against differences of the
  series themselves.
* The tracker (``andfn/dfn.py:2638-2834``): straight lines of step ds r in uniform flow, the disk
  clamp, exact crossings of known segments and circles, the even split of the exit at a symmetric
  intersection, and on golden networks the stream function constant along every traced segment.
"""

import ctypes

import numpy as np
import pytest

import golden_io
from oracle import oracle
from oracle import streamlines as osl


def _converged(name):
    g = golden_io.load(name)
    fa, ea, consts, K = golden_io.records(g)
    nc = g["out_coef"].shape[1]
    ea["coef"][:] = 0
    ea["coef"][:, :nc] = g["out_coef"]
    ea["ncoef"] = g["out_ncoef"]
    ea["nint"] = g["out_nint"]
    ea["q"] = g["out_q"]
    fa["constant"] = g["out_constant"]
    return fa, ea, K


def _seg_dist(z, a, b):
    t = np.clip(np.real((z - a) * np.conj(b - a)) / np.abs(b - a) ** 2, 0, 1)
    return np.abs(z - (a + t * (b - a)))


def _clearance(fa, ea, f, z):
    """Distance of points z of fracture f to its nearest element."""
    d = np.full(len(z), np.inf)
    for e in fa["elements"][f][: fa["nelements"][f]]:
        r = ea[e]
        ty = r["_type"]
        if ty == 1:
            d = np.minimum(d, r["radius"] - np.abs(z))
        elif ty in (2, 4):
            d = np.minimum(d, np.abs(np.abs(z - r["center"]) - r["radius"]))
        else:
            ep = r["endpoints1"] if (ty == 0 and r["frac0"] != f) else r["endpoints0"]
            d = np.minimum(d, _seg_dist(z, ep[0], ep[1]))
    return d


def _points(fa, ea, fracs=None, per_frac=30, seed=0, clearance=0.05):
    rng = np.random.default_rng(seed)
    frac, z = [], []
    for f in range(len(fa)) if fracs is None else fracs:
        R = fa["radius"][f]
        zz = R * np.sqrt(rng.uniform(0, 0.85, 4 * per_frac)) * np.exp(2j * np.pi * rng.uniform(size=4 * per_frac))
        zz = zz[_clearance(fa, ea, f, zz) > clearance * R][:per_frac]
        frac.append(np.full(len(zz), f))
        z.append(zz)
    return np.concatenate(frac), np.concatenate(z)


def _dOmega_dz(O, frac, z, h):
    """dOmega/dz = Phi_x - i Phi_y by central differences of Phi = Re Omega (single valued)."""
    phi = lambda zz: O.omega_at(frac, zz).real  # noqa: E731
    px = (phi(z + h) - phi(z - h)) / (2 * h)
    py = (phi(z + 1j * h) - phi(z - 1j * h)) / (2 * h)
    return px - 1j * py


@pytest.mark.parametrize("name", ["template_imp", "dfn40w"])
def test_object_form_is_minus_dOmega_dz(name):
    fa, ea, K = _converged(name)
    frac, z = _points(fa, ea)
    O = oracle.OracleHandle(fa, ea, K)
    try:
        w = O.w_at(frac, z, jit=False)
        fd = _dOmega_dz(O, frac, z, 1e-5 * fa["radius"][frac])
    finally:
        O.close()
    ok = np.isfinite(w) & np.isfinite(fd)
    assert ok.sum() > 0.9 * len(z)
    scale = np.max(np.abs(w[ok]))
    assert np.max(np.abs(w[ok] + fd[ok])) <= 1e-6 * scale


def _only(fa, ea, keep):
    """Records reduced to element `keep` (every other element's coefficients and q zeroed)."""
    ea = ea.copy()
    mask = np.arange(len(ea)) != keep
    ea["coef"][mask] = 0
    ea["q"][mask] = 0
    return fa.copy(), ea


@pytest.mark.parametrize("etype", [0, 1, 2, 3, 4, 5])
def test_jit_form_against_the_object_form(etype):
    name = "template_imp" if etype in (1, 4, 5) else "dfn40w"
    fa0, ea0, K = _converged(name)
    cand = np.nonzero(ea0["_type"] == etype)[0]
    assert len(cand), f"{name} has no element of type {etype}"
    e = int(cand[np.argmax(np.abs(ea0["coef"][cand]).sum(axis=1) + np.abs(ea0["q"][cand]))])
    fa, ea = _only(fa0, ea0, e)
    f = int(ea["frac0"][e])
    frac, z = _points(fa, ea, fracs=[f], per_frac=40, seed=3)
    O = oracle.OracleHandle(fa, ea, K)
    try:
        wj = O.w_at(frac, z, jit=True)
        wo = O.w_at(frac, z, jit=False)
        fd = _dOmega_dz(O, frac, z, 1e-5 * fa["radius"][f])
    finally:
        O.close()
    ok = np.isfinite(wo)
    scale = np.max(np.abs(fd[ok]))
    assert scale > 0
    assert np.max(np.abs(wo[ok] + fd[ok])) <= 1e-6 * scale
    if etype == 1:  # r dOmega/dz, no sign change
        assert np.max(np.abs(wj - fa["radius"][f] * fd)) <= 1e-6 * fa["radius"][f] * scale
    elif etype in (4, 5):  # not summed by hpc_fracture.calc_w
        assert np.all(wj == 0)
    else:
        assert np.max(np.abs(wj[ok] - wo[ok])) <= 1e-12 * scale


def _series_lib():
    lib = oracle.load()
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    for n in ("oracle_t_asym", "oracle_t_taylor", "oracle_t_asym_d1", "oracle_t_taylor_d1"):
        getattr(lib, n).argtypes = [i64, vp, vp, i64, vp]
    return lib


def _series(name, chi, coef):
    out = np.zeros_like(chi)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    getattr(_series_lib(), name)(len(chi), P(chi), P(coef), len(coef), P(out))
    return out


def test_series_derivatives():
    rng = np.random.default_rng(5)
    coef = rng.normal(size=12) + 1j * rng.normal(size=12)
    coef /= np.arange(1, 13) ** 1.5
    chi = 1.3 * np.exp(2j * np.pi * rng.uniform(size=50)) * rng.uniform(1, 2, 50)
    h = 1e-6
    d = (_series("oracle_t_asym", chi + h, coef) - _series("oracle_t_asym", chi - h, coef)) / (2 * h)
    assert np.allclose(_series("oracle_t_asym_d1", chi, coef), d, rtol=0, atol=1e-8)
    u = 0.8 * np.exp(2j * np.pi * rng.uniform(size=50)) * rng.uniform(0.2, 1, 50)
    d = (_series("oracle_t_taylor", u + h, coef) - _series("oracle_t_taylor", u - h, coef)) / (2 * h)
    assert np.allclose(_series("oracle_t_taylor_d1", u, coef), d, rtol=0, atol=1e-8)
    # a constant series has no derivative (the reference would index past the slice)
    assert np.all(_series("oracle_t_taylor_d1", u, coef[:1]) == 0)


def _uniform_flow(R=2.0, c1=0.3 - 0.4j, K=5):
    """One fracture with only a bounding circle: Omega = C + c1 z / R, w = -c1 / R everywhere."""
    fa, ea, _, _ = golden_io.records(golden_io.load("template"))
    keep = int(np.nonzero(ea["_type"] == 1)[0][0])
    f = int(ea["frac0"][keep])
    fa = fa[[f]].copy()
    ea = ea[[keep]].copy()
    ea["_id"] = 0
    ea["frac0"] = 0
    ea["radius"] = R
    ea["coef"] = 0
    ea["ncoef"] = 2
    ea["coef"][0, 1] = c1
    ea["q"] = 0
    fa["_id"] = 0
    fa["radius"] = R
    fa["elements"] = 0
    fa["nelements"] = 1
    return fa, ea, K


def test_tracker_uniform_flow_and_disk_clamp():
    fa, ea, K = _uniform_flow()
    O = oracle.OracleHandle(fa, ea, K)
    try:
        T = osl.Tracker(O, fa, ea, [1e-3])
        w = T.calc_w(0.1 + 0.2j, 0)
        assert abs(w - (-(0.3 - 0.4j) / 2.0)) < 1e-15
        z1 = T.runge_kutta(0.1 + 0.2j, 0, 0.02, False)
        step = np.conj(w) / abs(w) * 0.02
        assert abs(z1 - (0.1 + 0.2j + step)) < 1e-15
        assert abs(T.runge_kutta(0.1 + 0.2j, 0, 0.02, True) - (0.1 + 0.2j - step)) < 1e-15
        # a step that leaves the disk is pulled back just inside it
        zb = 1.99 * np.conj(w) / abs(w)
        z2 = T.runge_kutta(complex(zb), 0, 0.05, False)
        assert abs(z2) < 2.0 and abs(abs(z2) - 2.0 / (1 + 1e-5)) < 1e-12
        s, fr, v, e = T.track(0.0j, 0, 0.5, 0.01, 50, False)
        assert e == -1 and fr == [0] and len(s[0]) == 52  # the budget: 50 + 2 points
        pts = np.array(s[0])
        assert np.allclose(np.diff(pts), step, rtol=0, atol=1e-14)
        assert np.allclose(v[0], abs(w) / 1e-3, rtol=1e-14)
        t, le = osl.length_time(s[0], v[0])
        assert abs(le - np.sum(np.abs(np.diff(pts)))) < 1e-12 and abs(t - le / (abs(w) / 1e-3)) < 1e-9
    finally:
        O.close()


def test_crossings_of_known_segments_and_circles():
    ep = [-1.0 + 0.0j, 1.0 + 0.0j]
    z = osl.check_line_crossing(0.2 - 0.5j, 0.4 + 0.5j, ep)
    assert abs(z - 0.3) < 1e-15
    assert osl.check_line_crossing(0.2 - 0.5j, 0.4 - 0.1j, ep) is None  # step ends short
    assert osl.check_line_crossing(1.2 - 0.5j, 1.4 + 0.5j, ep) is None  # beyond the endpoint
    assert osl.check_line_crossing(0.2 + 0.5j, 0.6 + 0.5j, ep) is None  # parallel
    # a step into a well of radius 0.5 at 1+1j, entering along the real direction
    zc = osl.check_well_crossing(0.0 + 1.0j, 0.8 + 1.0j, 1 + 1j, 0.5)
    assert abs(zc - (0.5 * (1 + 1e-10) * -1 + 1 + 1j)) < 1e-12
    assert osl.check_well_crossing(0.0 + 0.0j, 0.3 + 0.0j, 1 + 1j, 0.5) is None


def _unwrapped_spread(psi, jumps, tol):
    """Spread of psi along a line once steps equal to a branch-cut jump (+-q of an element with a
    log term: the stream function of q/(2 pi) log chi jumps by q across its cut) are removed."""
    d = np.diff(psi)
    for k in range(len(d)):
        if abs(d[k]) > tol and len(jumps) and np.min(np.abs(np.abs(d[k]) - jumps)) <= tol:
            d[k] = 0.0
    c = np.concatenate([[0.0], np.cumsum(d)])
    return np.ptp(c)


@pytest.mark.parametrize("name", ["dfn40w", "dfn40"])
def test_stream_function_constant_along_segments(name):
    fa, ea, K = _converged(name)
    O = oracle.OracleHandle(fa, ea, K)
    try:
        T = osl.Tracker(O, fa, ea, np.full(len(fa), 1e-4))
        rng = np.random.default_rng(11)
        nseg = 0
        for f in rng.choice(len(fa), 4, replace=False):
            z0 = complex(0.3 * fa["radius"][f] * np.exp(2j * np.pi * rng.uniform()))
            s, fr, v, e = T.track(z0, int(f), 0.5, 0.01, 120, False)
            for seg, ff in zip(s, fr):
                seg = np.array(seg[1:-1])  # interior steps (the ends sit on elements)
                if len(seg) < 3:
                    continue
                psi = O.omega_at(np.full(len(seg), ff), seg).imag
                ring = fa["radius"][ff] * 0.8 * np.exp(2j * np.pi * np.arange(256) / 256)
                spread = np.ptp(O.omega_at(np.full(256, ff), ring).imag)
                ids = fa["elements"][ff][: fa["nelements"][ff]]
                jumps = np.abs(ea["q"][ids][ea["q"][ids] != 0])
                ok = np.isfinite(psi)
                # the predictor-corrector's own error grows where the flow bends sharply (next to
                # the elements the line ends on): a line along grad Phi instead would change psi by
                # a good part of the spread
                assert _unwrapped_spread(psi[ok], jumps, 1e-4 * spread) <= 2e-2 * spread
                nseg += 1
        assert nseg > 0
    finally:
        O.close()
