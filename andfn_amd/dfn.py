"""DFN object model and the packing of a network into solver records.

This is the caller side of the solve path, kept API-compatible with AnDFN so that scripts
written against ``andfn.DFN`` run unchanged on the accelerated solve:

* ``Fracture``/``BoundingCircle``/``Intersection``/``ConstantHeadLine``/``Well``/
  ``ImpermeableCircle``/``ImpermeableLine`` keep the constructor signatures of
  ``andfn/fracture.py:17-80``, ``andfn/bounding.py:14-46``, ``andfn/intersection.py:14-56``,
  ``andfn/const_head.py:14-49``, ``andfn/well.py:14-43``, ``andfn/impermeable_object.py``.
* ``DFN.get_elements`` reproduces the element order (bounding, impermeable circle,
  impermeable line, constant head, well, intersection) and ids = positions
  (``andfn/dfn.py:995-1063``).
* ``DFN.consolidate_dfn(hpc=True)`` packs the records exactly like
  ``andfn/dfn.py:658-820``, except that the arrays are zero-initialised where AnDFN leaves
  ``np.empty`` memory (the solver relies on a zero coefficient tail).
* ``DFN.solve`` calls :func:`andfn_amd.hpc_solve.solve`, which has the signature of
  ``andfn.hpc.hpc_solve.solve`` (``andfn/hpc/hpc_solve.py:55-79``).
"""

import logging

import numpy as np

from . import geometry as geo
from .dtypes import (
    BOUNDING_CIRCLE,
    CONST_HEAD_LINE,
    IMP_CIRCLE,
    IMP_LINE,
    INTERSECTION,
    WELL,
    MAX_ELEMENTS,
    MAX_NCOEF,
    default_constants,
    dtype_constants,
    make_element_dtype,
    make_fracture_dtype,
)

logger = logging.getLogger("andfn")


class Element:
    _type = -1

    def __init__(self, label, frac0, ncoef=5, nint=10):
        self.label = label
        self._id = -1
        self.frac0 = frac0
        self.frac1 = None
        self.ncoef = ncoef
        self.nint = nint
        self.q = 0.0
        self.coef = np.zeros(ncoef, dtype=complex)
        self.error = 1.0
        self.radius = 0.0
        self.center = 0.0 + 0.0j
        self.head = 0.0
        self.phi = 0.0
        self.endpoints0 = np.zeros(2, dtype=complex)
        self.endpoints1 = np.zeros(2, dtype=complex)
        frac0.add_element(self)

    def set_id(self, _id):
        self._id = _id


class BoundingCircle(Element):
    _type = BOUNDING_CIRCLE

    def __init__(self, label, radius, frac0, ncoef=5, nint=10):
        super().__init__(label, frac0, ncoef, nint)
        self.radius = radius
        self.center = 0.0 + 0.0j


class Intersection(Element):
    _type = INTERSECTION

    def __init__(self, label, endpoints0, endpoints1, frac0, frac1, ncoef=5, nint=10):
        super().__init__(label, frac0, ncoef, nint)
        self.frac1 = frac1
        frac1.add_element(self)
        self.endpoints0 = np.asarray(endpoints0, dtype=complex)
        self.endpoints1 = np.asarray(endpoints1, dtype=complex)


class ConstantHeadLine(Element):
    _type = CONST_HEAD_LINE

    def __init__(self, label, endpoints0, head, frac0, ncoef=5, nint=10):
        super().__init__(label, frac0, ncoef, nint)
        self.endpoints0 = np.asarray(endpoints0, dtype=complex)
        self.head = head
        self.phi = frac0.phi_from_head(head)


class Well(Element):
    _type = WELL

    def __init__(self, label, radius, center, head, frac0):
        super().__init__(label, frac0)
        self.radius = radius
        self.center = complex(center)
        self.head = head
        self.phi = frac0.phi_from_head(head)


class ImpermeableCircle(Element):
    _type = IMP_CIRCLE

    def __init__(self, label, radius, center, frac0, ncoef=5, nint=10):
        super().__init__(label, frac0, ncoef, nint)
        self.radius = radius
        self.center = complex(center)


class ImpermeableLine(Element):
    _type = IMP_LINE

    def __init__(self, label, endpoints0, frac0, ncoef=5, nint=10):
        super().__init__(label, frac0, ncoef, nint)
        self.endpoints0 = np.asarray(endpoints0, dtype=complex)


class Fracture:
    def __init__(self, label, t, radius, center, normal, aperture=None, ncoef=5, nint=10, elements=None,
                 basis=None):
        self.label = label
        self._id = 0
        self.t = t
        self.aperture = aperture
        self.radius = radius
        self.center = np.asarray(center, dtype=np.float64)
        self._raw_normal = np.asarray(normal, dtype=np.float64)
        if basis is None:
            basis = geo.fracture_basis(np.asarray(normal, dtype=np.float64))
        # (normal, x_vector, y_vector); a builder of many fractures passes rows of one batched
        # fracture_basis, which gives the same values as the per-fracture call
        self.normal, self.x_vector, self.y_vector = basis
        self.elements = []
        self.constant = 0.0
        if elements is None:
            BoundingCircle(label, radius, self, ncoef, nint)

    def set_id(self, _id):
        self._id = _id

    def add_element(self, e):
        if e not in self.elements:
            self.elements.append(e)

    def phi_from_head(self, head):
        return head * self.t

    def head_from_phi(self, phi):
        return phi / self.t

    def delete_all_elements(self):
        for e in list(self.elements):
            for f in (e.frac0, e.frac1):
                if f is not None and f is not self and e in f.elements:
                    f.elements.remove(e)
        self.elements = []


class DFN:
    """Discrete fracture network with the accelerated AEM solve.

    Mirrors ``andfn.DFN`` (``andfn/dfn.py:318-358``) for the parts the solve needs.
    Solver constants are held in a 0-d ``dtype_constants`` record (``self.constants``)
    and can be changed through keyword arguments, like ``DFN.set_kwargs``.
    """

    def __init__(self, label, discharge_int=50, max_ncoef=MAX_NCOEF, max_elements=MAX_ELEMENTS, **kwargs):
        self.label = label
        self.discharge_int = discharge_int
        self.fractures = []
        self.elements = None
        self.ntype_element = np.zeros(6, dtype=np.int64)
        self.constants = default_constants()
        self.max_ncoef = max_ncoef
        self.max_elements = max_elements
        self.elements_struc_array_hpc = None
        self.fractures_struc_array_hpc = None
        self.elements_struc_array = None
        self.last_solve_info = None
        self._flow_net = None
        self._flow_key = None
        self.set_kwargs(**kwargs)

    # -- configuration ---------------------------------------------------------------------
    def set_kwargs(self, **kwargs):
        for key, value in kwargs.items():
            if key in dtype_constants.names:
                if key == "MAX_NCOEF" and value > self.max_ncoef:
                    raise ValueError(f"MAX_NCOEF cannot be greater than {self.max_ncoef}")
                if key == "MAX_ELEMENTS" and value > self.max_elements:
                    raise ValueError(f"MAX_ELEMENTS cannot be greater than {self.max_elements}")
                self.constants[key] = value
            else:
                setattr(self, key, value)

    # -- construction ----------------------------------------------------------------------
    def add_fracture(self, new_fracture):
        if isinstance(new_fracture, (list, tuple)):
            self.fractures.extend(new_fracture)
        else:
            self.fractures.append(new_fracture)
        self.elements = None
        for i, f in enumerate(self.fractures):
            f.set_id(i)

    def delete_fracture(self, fracture):
        if isinstance(fracture, Fracture):
            fracture = [fracture]
        drop = set(id(f) for f in fracture)
        for f in fracture:
            f.delete_all_elements()
        self.fractures = [f for f in self.fractures if id(f) not in drop]
        self.elements = None
        for i, f in enumerate(self.fractures):
            f.set_id(i)

    @property
    def number_of_fractures(self):
        return len(self.fractures)

    def number_of_elements(self):
        if self.elements is None:
            self.get_elements()
        return len(self.elements)

    def _frac_arrays(self, fracs):
        c = np.array([f.center for f in fracs], dtype=np.float64).reshape(-1, 3)
        n = np.array([f.normal for f in fracs], dtype=np.float64).reshape(-1, 3)
        x = np.array([f.x_vector for f in fracs], dtype=np.float64).reshape(-1, 3)
        y = np.array([f.y_vector for f in fracs], dtype=np.float64).reshape(-1, 3)
        r = np.array([f.radius for f in fracs], dtype=np.float64)
        return c, n, x, y, r

    def get_fracture_intersections(self, ncoef=5, nint=10, se_factor=None, rim_clearance=0.0, device=False):
        """Add an Intersection for every intersecting pair i < k, in (i, k) order.

        Same pairs and creation order as ``andfn/dfn.py:1437-1455``; the O(N^2) pair loop is
        replaced by a KD-tree pre-filter on |c_i - c_k| <= r_i + r_k (+2e-8, the candidate
        tolerance), which cannot drop an intersecting pair.  ``rim_clearance`` > 0 (not in
        AnDFN; default off) skips an intersection whose shortened endpoints come closer than
        rim_clearance * R to the rim of either fracture: the segment is clipped to the part
        inside (1 - rim_clearance) R of both fractures, and dropped only when nothing remains.
        ``device`` runs the pair search and the chords on the GPU (``construction.py``; the same
        pairs and chords).
        """
        if se_factor is None:
            se_factor = float(self.constants["SE_FACTOR"])
        fr = self.fractures
        c, n, x, y, r = self._frac_arrays(fr)
        if device:
            from . import construction

            ds = construction.DiskSet(c, n, x, y, r)
            pairs = construction.candidate_pairs(ds, 1e-8)
            if len(pairs):
                i, k = pairs[:, 0], pairs[:, 1]
                e0, e1, ok = construction.chords(i, k, ds, ds, 3, se_factor, 1 if rim_clearance > 0 else 0,
                                                 1.0 - rim_clearance)
                for p in np.nonzero(ok)[0]:
                    a, b = fr[i[p]], fr[k[p]]
                    Intersection(f"{a.label}_{b.label}", e0[p], e1[p], a, b, ncoef, nint)
            self.get_elements()
            return
        pairs = geo.candidate_pairs(c, r + 1e-8)
        if len(pairs):
            i, k = pairs[:, 0], pairs[:, 1]
            e0, e1, ok = geo.fracture_intersection(c[i], n[i], x[i], y[i], r[i], c[k], n[k], x[k], y[k], r[k])
            e0 = geo.shorten_line(e0, se_factor)
            e1 = geo.shorten_line(e1, se_factor)
            if rim_clearance > 0:
                e0, e1, ok2 = geo.clip_to_disks(e0, r[i], e1, r[k], 1.0 - rim_clearance)
                ok = ok & ok2
            for p in np.nonzero(ok)[0]:
                a, b = fr[i[p]], fr[k[p]]
                Intersection(f"{a.label}_{b.label}", e0[p], e1[p], a, b, ncoef, nint)
        self.get_elements()

    def set_constant_head_boundary(self, center, normal, radius, head, label="Constant Head Boundary",
                                   ncoef=5, nint=10, se_factor=None, tolerance=-1, rim_clearance=0.0, device=False):
        """Constant-head lines where a boundary disk cuts the fractures (``andfn/geometry_functions.py:878-926``).

        ``rim_clearance`` > 0 (not in AnDFN; default off) clips a line to the part inside
        (1 - rim_clearance) R and skips it when nothing remains.
        """
        if se_factor is None:
            se_factor = float(self.constants["SE_FACTOR"])
        fr = self.fractures
        if not fr:
            return
        c, n, x, y, r = self._frac_arrays(fr)
        bc = np.asarray(center, dtype=np.float64)
        bn, bx, by = geo.fracture_basis(np.asarray(normal, dtype=np.float64))
        near = np.sum((bc - c) ** 2, axis=1) <= (radius + r) ** 2
        idx = np.nonzero(near)[0]
        if len(idx) == 0:
            return
        P = len(idx)
        if device:
            # the boundary disk against every nearby fracture on the GPU: shortened and clipped in
            # the fracture's plane (the constant-head line), unshortened in the boundary's
            from . import construction

            bd = construction.DiskSet(bc, bn, bx, by, [float(radius)])
            fd = construction.DiskSet(c, n, x, y, r)
            e_full = construction.chords(np.zeros(P, dtype=np.int64), idx, bd, fd, 0)[1]
            _, ends_d, ok = construction.chords(np.zeros(P, dtype=np.int64), idx, bd, fd, 2, se_factor,
                                                2 if rim_clearance > 0 else 0, 1.0 - rim_clearance)
            for p in np.nonzero(ok)[0]:
                f = fr[idx[p]]
                if abs(e_full[p][0] - e_full[p][1]) < tolerance * f.radius:
                    continue
                ConstantHeadLine(f"{label}_{f.label}", ends_d[p], head, f, ncoef, nint)
            return
        rep = lambda v: np.repeat(v[None, :], P, axis=0)
        e0, e1, ok = geo.fracture_intersection(rep(bc), rep(bn), rep(bx), rep(by), np.full(P, float(radius)),
                                               c[idx], n[idx], x[idx], y[idx], r[idx])
        for p in np.nonzero(ok)[0]:
            f = fr[idx[p]]
            length = abs(e1[p][0] - e1[p][1])
            if length < tolerance * f.radius:
                continue
            ends = geo.shorten_line(e1[p], se_factor)
            if rim_clearance > 0:
                cl, _, okc = geo.clip_to_disks(ends[None, :], np.array([f.radius]), rho=1.0 - rim_clearance)
                if not okc[0]:
                    continue
                ends = cl[0]
            ConstantHeadLine(f"{label}_{f.label}", ends, head, f, ncoef, nint)

    def fracture_graph(self):
        """Fracture adjacency through intersections and the boundary flag per fracture."""
        nf = len(self.fractures)
        idx = {id(f): i for i, f in enumerate(self.fractures)}
        a, b = [], []
        boundary = np.zeros(nf, dtype=bool)
        for i, f in enumerate(self.fractures):
            for e in f.elements:
                if e._type == INTERSECTION and e.frac0 is f:
                    a.append(i)
                    b.append(idx[id(e.frac1)])
                elif e._type in (WELL, CONST_HEAD_LINE):
                    boundary[i] = True
        return np.array(a, dtype=np.int64), np.array(b, dtype=np.int64), boundary

    def keep_largest_connected_component(self, device=False):
        """Delete every fracture outside the largest connected component that touches a boundary.

        A component without a constant-head line or well leaves its fracture constants
        undetermined (a singular discharge system), so only components with a boundary
        element are candidates.  AnDFN's ``check_connectivity`` (``andfn/dfn.py:1674-1695``)
        removes fractures not connected to any boundary; this keeps one component of those.
        """
        nf = len(self.fractures)
        a, b, boundary = self.fracture_graph()
        if device:
            from . import construction

            lab = construction.components(nf, a, b)  # the library's union-find (host C++)
            ncomp = int(lab.max()) + 1 if nf else 0
        else:
            from scipy.sparse import coo_matrix
            from scipy.sparse.csgraph import connected_components

            g = coo_matrix((np.ones(len(a)), (a, b)), shape=(nf, nf))
            ncomp, lab = connected_components(g, directed=False)
        has_bnd = np.zeros(ncomp, dtype=bool)
        has_bnd[lab[boundary]] = True
        sizes = np.bincount(lab, minlength=ncomp)
        sizes = np.where(has_bnd, sizes, -1)
        best = int(np.argmax(sizes))
        if sizes[best] <= 0:
            raise RuntimeError("no fracture is connected to a boundary condition")
        drop = [f for i, f in enumerate(self.fractures) if lab[i] != best]
        if drop:
            self.delete_fracture(drop)
        self.get_elements()

    def get_elements(self):
        seen = set()
        elements = []
        for f in self.fractures:
            for e in f.elements:
                if id(e) not in seen:
                    seen.add(id(e))
                    elements.append(e)
        by_type = {t: [] for t in range(6)}
        for e in elements:
            by_type[e._type].append(e)
        self.elements = (by_type[BOUNDING_CIRCLE] + by_type[IMP_CIRCLE] + by_type[IMP_LINE]
                         + by_type[CONST_HEAD_LINE] + by_type[WELL] + by_type[INTERSECTION])
        self.ntype_element = np.array([len(by_type[t]) for t in range(6)], dtype=np.int64)
        for i, e in enumerate(self.elements):
            e.set_id(i)

    # -- packing ---------------------------------------------------------------------------
    def consolidate_dfn(self, hpc=True):
        """Pack fractures and elements into the solver's records (zero-initialised)."""
        if self.elements is None:
            self.get_elements()
        edt = make_element_dtype(self.max_ncoef)
        fdt = make_fracture_dtype(self.max_elements)
        els = self.elements
        ea = np.zeros(len(els), dtype=edt)
        ea["_id"] = [e._id for e in els]
        ea["_type"] = [e._type for e in els]
        ea["frac0"] = [e.frac0._id for e in els]
        ea["frac1"] = [e.frac1._id if e.frac1 is not None else -1 for e in els]
        ea["radius"] = [e.radius for e in els]
        ea["center"] = [e.center for e in els]
        ea["head"] = [e.head for e in els]
        ea["phi"] = [e.phi for e in els]
        ea["q"] = [e.q for e in els]
        ea["ncoef"] = [e.ncoef for e in els]
        ea["nint"] = [e.nint for e in els]
        ea["error"] = [e.error for e in els]
        ea["endpoints0"] = np.array([e.endpoints0 for e in els], dtype=complex).reshape(-1, 2)
        ea["endpoints1"] = np.array([e.endpoints1 for e in els], dtype=complex).reshape(-1, 2)
        for i, e in enumerate(els):
            if e.ncoef > self.max_ncoef:
                raise IndexError(f"element {i}: ncoef {e.ncoef} exceeds capacity {self.max_ncoef}")
            nc = min(len(e.coef), e.ncoef)
            ea["coef"][i, :nc] = e.coef[:nc]
        fa = np.zeros(len(self.fractures), dtype=fdt)
        for i, f in enumerate(self.fractures):
            if len(f.elements) > self.max_elements:
                raise IndexError(f"fracture {i}: {len(f.elements)} elements exceed capacity {self.max_elements}")
            fa[i]["_id"] = f._id
            fa[i]["t"] = f.t
            fa[i]["radius"] = f.radius
            fa[i]["center"] = f.center
            fa[i]["normal"] = f.normal
            fa[i]["x_vector"] = f.x_vector
            fa[i]["y_vector"] = f.y_vector
            fa[i]["elements"][: len(f.elements)] = [e._id for e in f.elements]
            fa[i]["nelements"] = len(f.elements)
            fa[i]["constant"] = f.constant
        self.elements_struc_array_hpc = ea
        self.fractures_struc_array_hpc = fa
        self._flow_net = None

    def unconsolidate_dfn(self, hpc=True):
        ea = self.elements_struc_array_hpc
        for f, rec in zip(self.fractures, self.fractures_struc_array_hpc):
            f.constant = float(rec["constant"])
        for e, rec in zip(self.elements, ea):
            nc = int(rec["ncoef"])
            e.ncoef = nc
            e.nint = int(rec["nint"])
            e.coef = rec["coef"][:nc].copy()
            e.q = float(rec["q"])
            e.error = float(rec["error"])

    # -- checkpoint ------------------------------------------------------------------------
    def save_dfn(self, filename, overwrite=False):
        """Save the network and its solution (``andfn/dfn.py:404-458``); see ``checkpoint.py``."""
        from .checkpoint import save_dfn

        return save_dfn(self, filename, overwrite)

    def load_dfn(self, filename):
        """Load a saved network with its solution (``andfn/dfn.py:460-609``); the next solve
        starts from the saved coefficients and discharges."""
        from .checkpoint import load_dfn

        return load_dfn(self, filename)

    # -- solve -----------------------------------------------------------------------------
    def solve(self, unconsolidate=False, **options):
        """Solve the DFN on the GPU (drop-in for ``andfn/dfn.py:1807-1849``)."""
        from .hpc_solve import solve as hpc_solve

        if self.elements is None:
            self.get_elements()
        self.consolidate_dfn(hpc=True)
        self.elements_struc_array = hpc_solve(
            self.fractures_struc_array_hpc,
            self.elements_struc_array_hpc,
            self.discharge_int,
            self.constants,
            self.ntype_element,
            **options,
        )
        from .hpc_solve import last_solve_info

        self.last_solve_info = last_solve_info()
        self._flow_net = None
        if unconsolidate:
            self.unconsolidate_dfn(hpc=True)

    def get_dfn_discharge(self):
        """Half the summed |q| of the boundary elements (``andfn/dfn.py:1083-1092``)."""
        ea = self.elements_struc_array_hpc
        sel = (ea["_type"] == WELL) | (ea["_type"] == CONST_HEAD_LINE)
        return float(np.sum(np.abs(ea["q"][sel])) / 2)

    # -- flow on the solved network ----------------------------------------------------------
    def _device_network(self):
        """The solved records resident on the GPU (created on first use after a solve)."""
        from .hpc_solve import DeviceNetwork

        fa, ea = self.fractures_struc_array_hpc, self.elements_struc_array_hpc
        if fa is None or ea is None:
            raise RuntimeError("no solved network: call solve() first")
        if self._flow_net is None:
            self._flow_net = DeviceNetwork(fa, ea, self.discharge_int, self.constants)
        return self._flow_net

    def _frac_index(self, frac):
        return frac._id if isinstance(frac, Fracture) else int(frac)

    def calc_w(self, z, frac, exclude=None):
        """Discharge vector w at points z of fracture ``frac`` (``Fracture.calc_w``,
        ``andfn/fracture.py:453-477``), every element but ``exclude``."""
        zz = np.atleast_1d(np.asarray(z, dtype=np.complex128))
        f = np.full(len(zz), self._frac_index(frac), dtype=np.int64)
        ex = -1 if exclude is None else (exclude._id if isinstance(exclude, Element) else int(exclude))
        w = self._device_network().state_w(f, zz, exclude=ex, form="object")
        return w[0] if np.isscalar(z) else w

    def _apertures(self):
        ap = [f.aperture for f in self.fractures]
        if any(a is None for a in ap):
            raise ValueError("streamline speeds need an aperture on every fracture (|w| / aperture)")
        return np.array(ap, dtype=np.float64)

    def streamline_tracking(self, z0, frac, elevation, ds, max_length, backward):
        """Track one streamline from z0 on ``frac`` (``andfn/dfn.py:2638-2725``).

        Returns (streamline, streamline_frac, velocity, element) as the reference does: lists of
        per-fracture point and speed lists, the fractures visited, and the element the line left
        through (``False`` if none)."""
        s, f, v, e = self.track_streamlines(np.array([z0], dtype=np.complex128), frac, ds=ds, max_length=max_length,
                                            elevation=np.array([float(elevation)]), backward=backward)
        return s[0], f[0], v[0], e[0]

    def track_streamlines(self, z0, frac, ds=1e-2, max_length=1000, elevation=0.0, remove_false=True, backward=False):
        """Every streamline of ``plot_streamline_tracking`` (``andfn/dfn.py:2558-2636``: each start
        point with each elevation, start-point major), traced together on the GPU; returns
        (streamlines, streamlines_frac, velocities, elements) like that function, without the plot.
        ``remove_false`` only concerns the plot there and is accepted for the same signature."""
        from . import streamlines as sl

        z0 = np.atleast_1d(np.asarray(z0, dtype=np.complex128))
        elev = np.atleast_1d(np.asarray(elevation, dtype=np.float64))
        f = self._frac_index(frac)
        zz = np.repeat(z0, len(elev))
        ee = np.tile(elev, len(z0))
        if self.elements is None:
            self.get_elements()
        segs, segf, segv, el = sl.track(self._device_network(), self.fractures_struc_array_hpc,
                                        self.elements_struc_array_hpc, zz, f, ee, ds, max_length, backward,
                                        self._apertures())
        streamlines = [[p.tolist() for p in s] for s in segs]
        fracs = [[self.fractures[i] for i in fs] for fs in segf]
        vels = [[v.tolist() for v in vs] for vs in segv]
        elements = [self.elements[e] if e >= 0 else False for e in el]
        return streamlines, fracs, vels, elements

    @staticmethod
    def get_travel_time_and_length(streamline, velocity):
        """Travel time and length of a streamline (``andfn/dfn.py:2836-2883``)."""
        from .streamlines import get_travel_time_and_length

        return get_travel_time_and_length(streamline, velocity)

