"""CPU restatement of AnDFN's streamline tracking (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module; the product path never does.  It follows ``DFN.streamline_tracking``
(``andfn/dfn.py:2638-2725``) one point at a time in plain Python over the solver records, with the
discharge vector of the C++ oracle (``OracleHandle.w_at`` in the object model's form,
``andfn/fracture.py:453-477``):

* ``runge_kutta``              dfn.py:2779-2834
* ``check_streamline_exit``    dfn.py:2727-2742, over ``Intersection.check_chi_crossing``
  (intersection.py:227-271), ``ConstantHeadLine.check_chi_crossing`` (const_head.py:208-246) and
  ``Well.check_chi_crossing`` (well.py:162-200)
* ``line_line_intersection`` / ``line_circle_intersection``   geometry_functions.py:359-438
* ``get_exit_intersection``    dfn.py:2744-2777
* ``_get_length_time_fracture`` dfn.py:2886-2893

Parity of these functions is pinned by closed forms in tests/test_oracle_streamlines.py (uniform
flow, known crossings, the symmetric exit split) and by the stream function being constant along
traced lines.
"""

import cmath
import math

import numpy as np

INTERSECTION, WELL, CONST_HEAD_LINE = 0, 2, 3


def map_z_line_to_chi(z, ep):
    big_z = (2 * z - ep[0] - ep[1]) / (ep[1] - ep[0])
    return big_z + cmath.sqrt(big_z - 1) * cmath.sqrt(big_z + 1)


def map_chi_to_z_line(chi, ep):
    big_z = 0.5 * (chi + 1 / chi)
    return 0.5 * (big_z * (ep[1] - ep[0]) + ep[0] + ep[1])


def line_line_intersection(z0, z1, z2, z3):
    det = (z1.conjugate() - z0.conjugate()) * (z3 - z2) - (z1 - z0) * (z3.conjugate() - z2.conjugate())
    if det == 0:
        return None
    z = (z1.conjugate() * z0 - z1 * z0.conjugate()) * (z3 - z2) - (z1 - z0) * (z3.conjugate() * z2 - z3 * z2.conjugate())
    return z / det


def line_circle_intersection(z0, z1, radius):
    dx = (z1 - z0).real
    dy = (z1 - z0).imag
    if dx == 0:
        x = z0.real
        y1 = math.sqrt(radius**2 - x**2) if radius**2 - x**2 >= 0 else float("nan")
        return complex(x, y1), complex(x, -y1)
    m = dy / dx
    x0 = z0.imag - m * z0.real
    a = 1 + m**2
    b = 2 * x0 * m
    c = x0**2 - radius**2
    disc = b**2 - 4 * a * c
    if disc < 0:
        return None, None
    sd = math.sqrt(disc)
    x1 = (-b + sd) / (2 * a)
    x2 = (-b - sd) / (2 * a)
    return complex(x1, m * x1 + x0), complex(x2, m * x2 + x0)


def check_line_crossing(z0, z1, ep, atol=1e-12):
    z = line_line_intersection(z0, z1, ep[0], ep[1])
    if z is None:
        return None
    if abs(abs(z - z0) + abs(z1 - z) - abs(z0 - z1)) > atol:
        return None
    if abs(abs(z - ep[0]) + abs(z - ep[1]) - abs(ep[0] - ep[1])) > atol:
        return None
    return z


def check_well_crossing(z0, z1, center, radius, atol=1e-10):
    chi0 = (z0 - center) / radius
    chi1 = (z1 - center) / radius
    chi2, chi3 = line_circle_intersection(chi0, chi1, 1.0)
    if chi2 is None:
        return None
    diff0 = abs(abs(chi2 - chi0) + abs(chi2 - chi1) - abs(chi1 - chi0))
    diff1 = abs(abs(chi3 - chi0) + abs(chi3 - chi1) - abs(chi1 - chi0))
    if diff0 > atol and diff1 > atol:
        return None
    chi2 *= 1 + atol
    chi3 *= 1 + atol
    z2 = chi2 * radius + center
    z3 = chi3 * radius + center
    return z2 if abs(z2 - z0) < abs(z3 - z0) else z3


class Tracker:
    """Streamline tracking over solver records with the oracle's w (object-model form)."""

    def __init__(self, oracle_handle, fracture_struc_array, element_struc_array, aperture):
        self.O = oracle_handle
        self.fa = fracture_struc_array
        self.ea = element_struc_array
        self.ap = np.asarray(aperture, dtype=np.float64)

    def calc_w(self, z, f):
        return complex(self.O.w_at(np.array([f]), np.array([z], dtype=np.complex128), jit=False)[0])

    def velocity(self, z, f):
        return abs(self.calc_w(z, f)) / self.ap[f]

    def runge_kutta(self, z0, f, ds, backward, tolerance=1e-6, max_it=10):
        if backward:
            ds = -ds
        R = float(self.fa["radius"][f])
        w0 = self.calc_w(z0, f)
        if math.isnan(w0.real):
            return complex(float("nan"), float("nan"))
        z1 = z0 + w0.conjugate() / abs(w0) * ds
        if abs(z1) > R:
            z1 *= R / (abs(z1) * (1 + 1e-5))
        dz = 1e99
        it = 0
        while dz > tolerance and it < max_it:
            w1 = self.calc_w(z1, f)
            if math.isnan(w1.real):
                break
            z2 = z0 + (w0 + w1).conjugate() / abs(w0 + w1) * ds
            if abs(z2) > R:
                z2 *= R / (abs(z2) * (1 + 1e-5))
            dz = abs(z2 - z1)
            z1 = z2
            it += 1
        return z1

    def discharge_elements(self, f):
        ids = self.fa["elements"][f][: self.fa["nelements"][f]]
        return [int(e) for e in ids if self.ea["_type"][e] in (INTERSECTION, WELL, CONST_HEAD_LINE)]

    def check_exit(self, z0, z1, f):
        for e in self.discharge_elements(f):
            rec = self.ea[e]
            if rec["_type"] == WELL:
                z2 = check_well_crossing(z0, z1, complex(rec["center"]), float(rec["radius"]))
            else:
                second = rec["_type"] == INTERSECTION and rec["frac0"] != self.fa["_id"][f]
                ep = [complex(v) for v in (rec["endpoints1"] if second else rec["endpoints0"])]
                z2 = check_line_crossing(z0, z1, ep)
            if z2 is not None:
                return z2, e
        return None, -1

    def to_3d(self, z, f):
        return z.real * self.fa["x_vector"][f] + z.imag * self.fa["y_vector"][f] + self.fa["center"][f]

    def to_2d(self, p, f):
        d = p - self.fa["center"][f]
        return complex(np.dot(d, self.fa["x_vector"][f]), np.dot(d, self.fa["y_vector"][f]))

    def exit_intersection(self, z3d, e, f, f_old, elevation, dchi=1e-4):
        rec = self.ea[e]
        ep = [complex(v) for v in (rec["endpoints0"] if f == rec["frac0"] else rec["endpoints1"])]
        z = self.to_2d(z3d, f)
        chi0 = map_z_line_to_chi(z, ep)
        chi1 = chi0.conjugate()
        z0 = map_chi_to_z_line(chi0 * (1 + dchi), ep)
        z1 = map_chi_to_z_line(chi1 * (1 + dchi), ep)
        abs_w0 = abs(self.calc_w(z0, f))
        abs_w1 = abs(self.calc_w(z1, f))
        divide = abs_w0 / (abs_w0 + abs_w1) if (abs_w0 + abs_w1) != 0 else float("nan")
        nz0 = float(np.dot(self.to_3d(z0, f) - z3d, self.fa["normal"][f_old]))
        if nz0 < 0:
            up, down = z1, z0
        else:
            up, down = z0, z1
            divide = 1 - divide
        if elevation < divide:
            return down, z0, elevation / divide
        return up, z0, (elevation - divide) / (1 - divide)

    def track(self, z0, f, elevation, ds, max_length, backward):
        """(segments, fractures, speeds, exit element or -1), as DFN.streamline_tracking."""
        streamline, fracs, velocity = [], [], []
        element = -1
        z_start = z0
        while True:
            ds_frac = float(self.fa["radius"][f]) * ds
            length = sum(len(s) for s in streamline)
            psi = [z_start]
            w = [self.velocity(z0, f)]
            z1 = self.runge_kutta(z0, f, ds_frac, backward)
            if math.isnan(z1.real) or math.isnan(z1.imag):
                break
            z3, element = self.check_exit(z0, z1, f)
            while z3 is None:
                psi.append(z1)
                w.append(self.velocity(z1, f))
                z0 = z1
                z1 = self.runge_kutta(z0, f, ds_frac, backward)
                if math.isnan(z1.real) or math.isnan(z1.imag):
                    z3 = z0
                    break
                z3, element = self.check_exit(z0, z1, f)
                if len(psi) > max_length - length:
                    z3 = z1
                    break
            psi.append(z3)
            w.append(self.velocity(z3, f))
            streamline.append(psi)
            fracs.append(f)
            velocity.append(w)
            if element >= 0 and self.ea["_type"][element] == INTERSECTION:
                z3d = self.to_3d(z3, f)
                f_old = f
                rec = self.ea[element]
                f = int(rec["frac1"]) if f == rec["frac0"] else int(rec["frac0"])
                z0, z_start, elevation = self.exit_intersection(z3d, element, f, f_old, elevation)
            else:
                break
        return streamline, fracs, velocity, element


def length_time(streamline, velocity):
    """(time, length) of one segment (dfn.py:2886-2893)."""
    s = np.asarray(streamline)
    v = np.asarray(velocity)
    mask = ~np.isnan(s) & ~np.isnan(v)
    s, v = s[mask], v[mask]
    return float(np.sum(np.abs(s[1:] - s[:-1]) / v[:-1])), float(np.sum(np.abs(s[1:] - s[:-1])))
