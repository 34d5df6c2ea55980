"""Discharge vector and streamline tracking on the GPU against the oracle.

* ``get_w`` in both forms (``hpc_fracture.calc_w``, ``andfn/hpc/hpc_fracture.py:139-174``, and the
  object model's ``Fracture.calc_w``, ``andfn/fracture.py:453-477``) at points of every fracture of
  two golden converged states (every element type): to 1e-9 of the largest |w|, NaN at the same
  points.
* ``streamlines.track`` (segments on the device, exits through intersections on the host) against
  the oracle's restatement of ``DFN.streamline_tracking`` (``andfn/dfn.py:2638-2834``) on the same
  records: the same fractures visited, the same point counts and exit elements, every point to
  1e-8 of its fracture's radius and every speed to 1e-8 relative.  Two values are decided by
  rounding in the reference itself and are checked for what is fixed: the first point after a
  fracture change (one of two mirror images across the intersection) and the speed at an exit point
  on a line element (either side's).  Otherwise the tracking has no tolerance that rounding could
  cross except its 1e-6 corrector test and its 1e-12 crossing tests, which a difference in the last
  bits does not reach on these lines.
* ``DFN.streamline_tracking`` / ``track_streamlines`` / ``get_travel_time_and_length`` on a network
  solved on the GPU: the reference's return shapes and the oracle's lines.
"""

import numpy as np
import pytest

import golden_io
from andfn_amd import streamlines as sl
from andfn_amd.generator import build_network, workload
from andfn_amd.hpc_solve import DeviceNetwork, get_w
from oracle import oracle
from oracle import streamlines as osl

pytestmark = pytest.mark.gpu


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
    return fa, ea, consts, K


def _points(fa, per_frac=200, seed=0):
    rng = np.random.default_rng(seed)
    frac = np.repeat(np.arange(len(fa)), per_frac)
    r = fa["radius"][frac] * np.sqrt(rng.uniform(0, 0.98, len(frac)))
    return frac, r * np.exp(2j * np.pi * rng.uniform(size=len(frac)))


@pytest.mark.parametrize("name", ["template_imp", "dfn40w"])
def test_w_matches_oracle(name):
    print(f"w {name}", flush=True)
    fa, ea, _, K = _converged(name)
    frac, z = _points(fa)
    O = oracle.OracleHandle(fa.copy(), ea.copy(), K)
    try:
        for form, jit in (("jit", True), ("object", False)):
            ref = O.w_at(frac, z, jit=jit)
            got = get_w(fa, ea, frac, z, form=form)
            nan = np.isnan(ref)
            assert np.array_equal(np.isnan(got), nan)
            fin = ~nan & np.isfinite(ref)
            scale = np.max(np.abs(ref[fin]))
            err = np.max(np.abs(got[fin] - ref[fin])) / scale
            print(f"{name} {form}: {fin.sum()} points, max |w - oracle| {err:.2e} of {scale:.3e}")
            assert err < 1e-9
        # one element left out (the exclude argument of both calc_w forms)
        ex = int(fa["elements"][0][0])
        ref = O.w_at(frac, z, exclude=ex, jit=False)
        got = get_w(fa, ea, frac, z, exclude=ex, form="object")
        fin = np.isfinite(ref)
        assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-9 * np.max(np.abs(ref[fin]))
    finally:
        O.close()


def _mirrored(fa, ea, f, a, b, tol):
    """a and b are mirror images across one of fracture f's intersection lines."""
    for e in fa["elements"][f][: fa["nelements"][f]]:
        if ea["_type"][e] != 0:
            continue
        ep = ea["endpoints0"][e] if ea["frac0"][e] == f else ea["endpoints1"][e]
        c, h = 0.5 * (ep[0] + ep[1]), 0.5 * (ep[1] - ep[0])
        if abs(c + h * np.conj((a - c) / h) - b) <= tol:
            return True
    return False


def _speed_either_side(O, fa, ea, ap, f, z, v, rtol=1e-5):
    """v is the speed on one of the two sides of a line element of fracture f through z."""
    R = fa["radius"][f]
    for e in fa["elements"][f][: fa["nelements"][f]]:
        if ea["_type"][e] not in (0, 3):
            continue
        ep = ea["endpoints1"][e] if (ea["_type"][e] == 0 and ea["frac0"][e] != f) else ea["endpoints0"][e]
        d = ep[1] - ep[0]
        t = np.real((z - ep[0]) * np.conj(d)) / abs(d) ** 2
        if not (0 <= t <= 1 and abs(z - (ep[0] + t * d)) <= 1e-9 * R):
            continue
        n = 1j * d / abs(d) * 1e-7 * R
        w = O.w_at(np.array([f, f]), np.array([z + n, z - n]), jit=False)
        if np.min(np.abs(np.abs(w) / ap[f] - v)) <= rtol * v:
            return True
    return False


def _compare_lines(fa, ea, got, ref, label, O=None, ap=None):
    gs, gf, gv, ge = got
    rs, rf, rv, re = ref
    assert gf == rf, f"{label}: fractures {gf} vs {rf}"
    assert ge == re, f"{label}: exit element {ge} vs {re}"
    assert [len(s) for s in gs] == [len(s) for s in rs], label
    for k, (s, r, v, w, f) in enumerate(zip(gs, rs, gv, rv, gf)):
        R = fa["radius"][f]
        r = np.array(r)
        s = np.asarray(s)
        if k > 0 and abs(s[0] - r[0]) > 1e-8 * R:
            # The first point recorded after a fracture change is the candidate start on the chi
            # side of the intersection's map (get_exit_intersection, dfn.py:2756-2762, returns its
            # z0 whichever side the line continues on).  The exit point lies on the intersection
            # line, where the map's branch is picked by the sign of a rounding-level imaginary
            # part, so the reference itself may record either of the two mirror images; the side
            # the line continues on does not depend on it.
            assert _mirrored(fa, ea, f, s[0], r[0], 1e-8 * R), label
            s, r = s[1:], r[1:]
        assert np.max(np.abs(s - r)) <= 1e-8 * R, label
        w = np.array(w)
        v = np.asarray(v)
        if abs(v[-1] - w[-1]) > 1e-8 * abs(w[-1]):
            # An exit point through a line element lies on that line, where w jumps (a line sink
            # or source) and the side evaluated is picked by rounding: the reference's own speed
            # there is either side's.
            assert _speed_either_side(O, fa, ea, ap, f, np.asarray(gs[k])[-1], v[-1]), label
            v, w = v[:-1], w[:-1]
        fin = np.isfinite(w)
        assert np.array_equal(np.isfinite(v), fin), label
        assert np.max(np.abs(v[fin] - w[fin]) / np.abs(w[fin]), initial=0.0) <= 1e-8, label


@pytest.mark.parametrize("backward", [False, True])
def test_tracks_match_oracle(backward):
    print(f"tracks backward={backward}", flush=True)
    fa, ea, consts, K = _converged("dfn40w")
    ap = np.linspace(1e-4, 3e-4, len(fa))
    rng = np.random.default_rng(7)
    fr = rng.choice(len(fa), 12)
    z0 = 0.6 * fa["radius"][fr] * np.sqrt(rng.uniform(size=12)) * np.exp(2j * np.pi * rng.uniform(size=12))
    elev = rng.uniform(size=12)
    net = DeviceNetwork(fa, ea, K, consts)
    segs, segf, segv, el = sl.track(net, fa, ea, z0, fr, elev, 0.02, 150, backward, ap)
    O = oracle.OracleHandle(fa.copy(), ea.copy(), K)
    try:
        T = osl.Tracker(O, fa, ea, ap)
        crossings = 0
        for i in range(12):
            ref = T.track(complex(z0[i]), int(fr[i]), float(elev[i]), 0.02, 150, backward)
            _compare_lines(fa, ea, (segs[i], segf[i], segv[i], int(el[i])), ref, f"line {i}", O, ap)
            crossings += len(ref[1]) - 1
        print(f"12 lines, {crossings} fracture changes, {sum(len(s) for ss in segs for s in ss)} points")
    finally:
        O.close()


def test_dfn_streamline_api():
    print("DFN streamline API", flush=True)
    dfn = build_network(workload("dfn40w"))
    for f in dfn.fractures:
        f.aperture = 2e-4
    dfn.solve()
    fa, ea = dfn.fractures_struc_array_hpc, dfn.elements_struc_array_hpc
    frac = dfn.fractures[int(np.argmax(fa["nelements"]))]
    z0 = np.array([0.2 + 0.1j, -0.3 + 0.2j]) * frac.radius
    streamlines, fracs, vels, elements = dfn.track_streamlines(z0, frac, ds=0.02, max_length=100,
                                                               elevation=np.array([0.3, 0.7]))
    assert len(streamlines) == 4 and len(fracs) == 4 and len(vels) == 4 and len(elements) == 4
    O = oracle.OracleHandle(fa.copy(), ea.copy(), dfn.discharge_int)
    try:
        T = osl.Tracker(O, fa, ea, np.full(len(fa), 2e-4))
        for k, (zz, e) in enumerate([(z0[0], 0.3), (z0[0], 0.7), (z0[1], 0.3), (z0[1], 0.7)]):
            ref = T.track(complex(zz), frac._id, e, 0.02, 100, False)
            got = (streamlines[k], [f._id for f in fracs[k]], vels[k],
                   elements[k]._id if elements[k] is not False else -1)
            _compare_lines(fa, ea, got, ref, f"line {k}", O, np.full(len(fa), 2e-4))
            assert all(isinstance(p, list) for p in streamlines[k])
            t, le = dfn.get_travel_time_and_length(streamlines[k], vels[k])
            # the oracle's sums over the same points (the lines themselves are compared above)
            rt = sum(osl.length_time(s, v)[0] for s, v in zip(streamlines[k], vels[k]))
            rl = sum(osl.length_time(s, v)[1] for s, v in zip(streamlines[k], vels[k]))
            assert abs(t - rt) <= 1e-8 * rt and abs(le - rl) <= 1e-8 * rl
    finally:
        O.close()
    s, f, v, e = dfn.streamline_tracking(complex(z0[0]), frac, 0.3, 0.02, 100, False)
    assert s == streamlines[0] and [x._id for x in f] == [x._id for x in fracs[0]]
    w = dfn.calc_w(complex(z0[0]), frac)
    assert isinstance(w, complex) and abs(w) > 0
