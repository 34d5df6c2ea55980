"""Seeded synthetic DFN workloads.

A network is drawn from a counter-based generator (SplitMix64 of seed, stream and index),
so the same seed gives bit-identical inputs on any machine and numpy version:

* centres uniform in the cube [0, L]^3;
* normals uniform on the sphere (normalised iid standard normals, Box-Muller);
* radii from the truncated power law with density proportional to r^-a on [rmin, rmax]
  (inverse CDF);
* transmissivity T = 10^N(mu, sigma);
* constant-head boundary disks in the planes x = 0 (head h0) and x = L (head h1), large
  enough to cover every fracture that can reach the plane;
* the largest connected component that touches a boundary is kept;
* optional wells (fixed head) placed on kept fractures, clear of every other element.

Named workloads (``WORKLOADS``) fix the laws used by the tests and the benchmark.
"""

from dataclasses import dataclass, field

import numpy as np

from .dfn import DFN, Fracture, Well

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n):
    """n uniforms in [0, 1) for (seed, stream): splitmix64 of a 64-bit counter."""
    with np.errstate(over="ignore"):
        base = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream) * np.uint64(0x9E3779B1))
        ctr = base + np.arange(n, dtype=np.uint64)
        u = _splitmix64(ctr)
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """n standard normals by Box-Muller on two uniform streams."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1000003, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def power_law(u, rmin, rmax, a):
    """Inverse CDF of the density proportional to r^-a on [rmin, rmax] (a != 1)."""
    e = 1.0 - a
    return (rmin**e + u * (rmax**e - rmin**e)) ** (1.0 / e)


@dataclass
class NetworkSpec:
    n_fractures: int
    cube: float
    rmin: float
    rmax: float
    seed: int = 0
    alpha: float = 2.5
    log10_t_mean: float = -5.0
    log10_t_sd: float = 0.5
    ncoef: int = 10
    nint: int = 20
    head0: float = 1.0
    head1: float = 0.0
    n_wells: int = 0
    well_radius: float = 0.1
    well_head: float = 0.5
    discharge_int: int = 50
    rim_clearance: float = 0.0
    min_separation: float = 0.0  # drop near-coincident element pairs (see drop_near_coincident)
    min_angle_deg: float = 20.0
    constants: dict = field(default_factory=dict)


# stream ids
_S_CENTER, _S_NORMAL, _S_RADIUS, _S_T, _S_WELL = 1, 10, 20, 30, 40


def draw_fractures(spec):
    n = spec.n_fractures
    s = spec.seed
    centers = np.stack([uniform(s, _S_CENTER + j, n) * spec.cube for j in range(3)], axis=1)
    normals = np.stack([normal(s, _S_NORMAL + 2 * j, n) for j in range(3)], axis=1)
    radii = power_law(uniform(s, _S_RADIUS, n), spec.rmin, spec.rmax, spec.alpha)
    t = 10.0 ** (spec.log10_t_mean + spec.log10_t_sd * normal(s, _S_T, n))
    return centers, normals, radii, t


def _segment_distance(p, a, b):
    ab = b - a
    tt = np.clip(np.real((p - a) * np.conj(ab)) / max(abs(ab) ** 2, 1e-300), 0.0, 1.0)
    return abs(p - (a + tt * ab))


def _place_wells(dfn, spec):
    """Wells on kept fractures: a fracture drawn uniformly, a point uniform on 0.8 of its disk,
    rejected while closer than 2 well radii (+0.5 m) to any element of that fracture."""
    fr = dfn.fractures
    placed = 0
    k = 0
    while placed < spec.n_wells:
        u = uniform(spec.seed, _S_WELL + k, 3)
        k += 1
        f = fr[min(int(u[0] * len(fr)), len(fr) - 1)]
        rr = 0.8 * f.radius * np.sqrt(u[1])
        z = rr * np.exp(2j * np.pi * u[2])
        clear = 2 * spec.well_radius + 0.5
        ok = True
        for e in f.elements:
            if e._type in (0, 3, 5):
                ends = e.endpoints0 if (e._type != 0 or e.frac0 is f) else e.endpoints1
                if _segment_distance(z, ends[0], ends[1]) < clear:
                    ok = False
                    break
            elif e._type in (2, 4):
                if abs(z - e.center) < e.radius + clear:
                    ok = False
                    break
        if ok:
            Well(f"well_{placed}", spec.well_radius, z, spec.well_head, f)
            placed += 1
        if k > 100000:
            raise RuntimeError("could not place wells")


def _seg_dist(a, b, c, d):
    """Minimum distance between segments ab and cd (batched complex arrays); 0 when they cross."""
    def pt_seg(p, u, v):
        uv = v - u
        t = np.clip(np.real((p - u) * np.conj(uv)) / np.maximum(np.abs(uv) ** 2, 1e-300), 0.0, 1.0)
        return np.abs(p - (u + t * uv))

    dist = np.minimum(np.minimum(pt_seg(a, c, d), pt_seg(b, c, d)), np.minimum(pt_seg(c, a, b), pt_seg(d, a, b)))
    den = np.imag(np.conj(b - a) * (d - c))
    safe = np.where(np.abs(den) > 0, den, 1.0)
    s = np.imag(np.conj(c - a) * (d - c)) / safe
    t = np.imag(np.conj(c - a) * (b - a)) / safe
    cross = (np.abs(den) > 0) & (s > 0) & (s < 1) & (t > 0) & (t < 1)
    return np.where(cross, 0.0, dist)


def drop_near_coincident(dfn, rel_dist, min_angle_deg):
    """Delete line elements that nearly coincide with another line element of the same fracture.

    Two intersection / constant-head segments of one fracture closer than rel_dist times the
    shorter length and at an angle below min_angle_deg represent (almost) the same jump twice;
    AnDFN's element-by-element (Jacobi) iteration does not converge on such pairs.  The element
    created later is removed from both of its fractures.  Returns the number removed.
    """
    cos_min = np.cos(np.radians(min_angle_deg))
    drop = set()
    for f in dfn.fractures:
        lines = [e for e in f.elements if e._type in (0, 3) and id(e) not in drop]
        if len(lines) < 2:
            continue
        ends = np.array([(e.endpoints0 if (e._type != 0 or e.frac0 is f) else e.endpoints1) for e in lines])
        a, b = ends[:, 0], ends[:, 1]
        i, k = np.triu_indices(len(lines), 1)
        L = np.minimum(np.abs(b[i] - a[i]), np.abs(b[k] - a[k]))
        dist = _seg_dist(a[i], b[i], a[k], b[k])
        u = (b[i] - a[i]) / np.abs(b[i] - a[i])
        v = (b[k] - a[k]) / np.abs(b[k] - a[k])
        par = np.abs(np.real(u * np.conj(v))) > cos_min
        bad = (dist < rel_dist * L) & par
        for p in np.nonzero(bad)[0]:
            ei, ek = lines[i[p]], lines[k[p]]
            if id(ei) in drop or id(ek) in drop:
                continue
            drop.add(id(ek))
    removed = 0
    for f in dfn.fractures:
        keep = [e for e in f.elements if id(e) not in drop]
        removed += len(f.elements) - len(keep)
        f.elements = keep
    dfn.elements = None
    return removed


def build_network(spec, device=False):
    """Build and return the DFN for ``spec`` (elements collected, not yet consolidated).

    ``device`` runs the pair search, the intersection and constant-head chords on the GPU and the
    connectivity in the library's C++ (the same network).
    """
    from .geometry import fracture_basis

    centers, normals, radii, t = draw_fractures(spec)
    dfn = DFN(f"synthetic-{spec.n_fractures}-{spec.seed}", discharge_int=spec.discharge_int, **spec.constants)
    nb, xb, yb = fracture_basis(normals)
    fr = [Fracture(f"{i}", t[i], radii[i], centers[i], normals[i], ncoef=spec.ncoef, nint=spec.nint,
                   basis=(nb[i], xb[i], yb[i]))
          for i in range(spec.n_fractures)]
    dfn.add_fracture(fr)
    dfn.get_fracture_intersections(spec.ncoef, spec.nint, rim_clearance=spec.rim_clearance, device=device)
    L = spec.cube
    reach = np.sqrt(2.0) * (0.5 * L + spec.rmax) + spec.rmax
    dfn.set_constant_head_boundary(np.array([0.0, 0.5 * L, 0.5 * L]), np.array([1.0, 0.0, 0.0]), reach,
                                   spec.head0, label="x0", ncoef=spec.ncoef, nint=spec.nint,
                                   rim_clearance=spec.rim_clearance, device=device)
    dfn.set_constant_head_boundary(np.array([L, 0.5 * L, 0.5 * L]), np.array([1.0, 0.0, 0.0]), reach,
                                   spec.head1, label="xL", ncoef=spec.ncoef, nint=spec.nint,
                                   rim_clearance=spec.rim_clearance, device=device)
    if spec.min_separation > 0:
        drop_near_coincident(dfn, spec.min_separation, spec.min_angle_deg)
    dfn.keep_largest_connected_component(device=device)
    if spec.n_wells:
        _place_wells(dfn, spec)
    dfn.get_elements()
    return dfn


WORKLOADS = {
    # Benchmark workloads: the configured laws with fixed coefficient counts (MAX_NCOEF = ncoef, so
    # AnDFN's growth heuristic is off), every element clipped to 97% of its fracture radius, and
    # near-coincident element pairs thinned.  Without these two input filters AnDFN's damped Jacobi
    # iteration diverges on these networks: chords that graze a fracture rim leave endpoints within
    # 0.1% of the bounding circle, and bundles of nearly coplanar fractures put several almost
    # coincident intersections into one fracture.
    "cfg1": NetworkSpec(200, 100.0, 5.0, 50.0, ncoef=10, nint=20, rim_clearance=0.03, min_separation=0.05,
                        constants=dict(MAX_ERROR=1e-9, MAX_NCOEF=10, MAX_ITERATIONS=500)),
    "cfg2": NetworkSpec(5000, 500.0, 10.0, 150.0, ncoef=20, nint=40, rim_clearance=0.03, min_separation=0.05,
                        constants=dict(MAX_ERROR=1e-9, MAX_NCOEF=20, MAX_ITERATIONS=500)),
    "cfg3": NetworkSpec(20000, 1000.0, 10.0, 300.0, ncoef=20, nint=40, n_wells=50, rim_clearance=0.03,
                        min_separation=0.05, constants=dict(MAX_ERROR=1e-9, MAX_NCOEF=20, MAX_ITERATIONS=500)),
    # The same laws built exactly as AnDFN builds them (no clipping, AnDFN's default growth)
    # CPU-runnable FP64 case: 200 disks in a 100 m cube
    "dfn200": NetworkSpec(200, 100.0, 5.0, 50.0, ncoef=10, nint=20, constants=dict(MAX_ERROR=1e-9)),
    # FP64 GPU case: 5,000 disks in a 500 m cube
    "dfn5000": NetworkSpec(5000, 500.0, 10.0, 150.0, ncoef=20, nint=40, constants=dict(MAX_ERROR=1e-9)),
    # mixed-precision case: 20,000 disks in a 1,000 m cube with 50 wells
    "dfn20000": NetworkSpec(20000, 1000.0, 10.0, 300.0, ncoef=20, nint=40, n_wells=50,
                            constants=dict(MAX_ERROR=1e-9)),
    # small cases for tests
    "dfn40": NetworkSpec(40, 40.0, 5.0, 25.0, ncoef=10, nint=20, constants=dict(MAX_ERROR=1e-9)),
    "dfn40w": NetworkSpec(40, 40.0, 5.0, 25.0, seed=3, ncoef=10, nint=20, n_wells=2,
                          constants=dict(MAX_ERROR=1e-9)),
}


def workload(name, **override):
    """The named network spec with fields replaced by ``override``; a ``constants`` override is
    merged into the spec's own solver constants."""
    spec = WORKLOADS[name]
    if override:
        d = dict(spec.__dict__)
        if "constants" in override:
            override = dict(override)
            override["constants"] = {**spec.constants, **override["constants"]}
        d.update(override)
        spec = NetworkSpec(**d)
    return spec
