"""Many independent DFN realisations solved as one network.

Realisations share no fracture, so their union is one record set whose fracture graph has one
connected component per realisation.  The solver detects the components and keeps a separate
dense discharge block for each (factored by one strided-batched LU, solved by batched blocked
triangular kernels), while every element kernel (sweeps, fits, growth, adjoint) covers all
realisations in the same launch.  The solve stops when the largest head residual over all
realisations is at most the tolerance.

``concatenate`` offsets element ids, fracture links and per-fracture element lists; ``split``
returns per-realisation views of the combined arrays.  ``realisation_specs`` draws the batched
benchmark family: fracture counts uniform on [500, 4000], the 200-disk law (radii [5, 50] m, pdf
r^-2.5, log10 T ~ N(-5, 0.5)) in a cube scaled as (N / 200)^(1/3) x 100 m so that every
realisation has the density of the 200-disk network, and 10 to 20 coefficients per element;
elements are clipped to 95% of their fracture radius (with 97%, 4 of 64 realisations make
AnDFN's damped iteration diverge; with 95% all converge, in 44-76 iterations on those four).
"""

from dataclasses import replace

import numpy as np

from .generator import NetworkSpec, build_network, uniform

_S_BATCH_N, _S_BATCH_M = 90, 91


def concatenate(networks):
    """networks: list of (fracture_records, element_records).  Returns (fa, ea, offsets) with
    offsets[r] = (first fracture, first element) of realisation r (plus a final end entry)."""
    if not networks:
        raise ValueError("no networks")
    fdt, edt = networks[0][0].dtype, networks[0][1].dtype
    for fa, ea in networks:
        if fa.dtype != fdt or ea.dtype != edt:
            raise ValueError("all realisations need the same record capacities")
    nf = sum(len(fa) for fa, _ in networks)
    ne = sum(len(ea) for _, ea in networks)
    fa_all = np.zeros(nf, dtype=fdt)
    ea_all = np.zeros(ne, dtype=edt)
    offsets = []
    f0 = e0 = 0
    for fa, ea in networks:
        offsets.append((f0, e0))
        F, E = len(fa), len(ea)
        fa_all[f0:f0 + F] = fa
        ea_all[e0:e0 + E] = ea
        fa_all["_id"][f0:f0 + F] = fa["_id"] + f0
        for f in range(F):
            n = int(fa["nelements"][f])
            fa_all["elements"][f0 + f, :n] = fa["elements"][f, :n] + e0
        ea_all["_id"][e0:e0 + E] = ea["_id"] + e0
        ea_all["frac0"][e0:e0 + E] = ea["frac0"] + f0
        fr1 = ea["frac1"]
        ea_all["frac1"][e0:e0 + E] = np.where(fr1 >= 0, fr1 + f0, fr1)
        f0 += F
        e0 += E
    offsets.append((f0, e0))
    return fa_all, ea_all, offsets


def split(fa_all, ea_all, offsets):
    """Per-realisation views (fractures, elements) of combined records."""
    return [(fa_all[offsets[r][0]:offsets[r + 1][0]], ea_all[offsets[r][1]:offsets[r + 1][1]])
            for r in range(len(offsets) - 1)]


def realisation_specs(count=64, seed=2024, nmin=500, nmax=4000, cmin=10, cmax=20):
    """The batched family: seeded realisations of the 200-disk law at constant density."""
    un = uniform(seed, _S_BATCH_N, count)
    um = uniform(seed, _S_BATCH_M, count)
    specs = []
    for r in range(count):
        n = nmin + int(un[r] * (nmax - nmin + 1))
        m = cmin + int(um[r] * (cmax - cmin + 1))
        cube = 100.0 * (n / 200.0) ** (1.0 / 3.0)
        specs.append(NetworkSpec(n, cube, 5.0, 50.0, seed=seed + r, ncoef=m, nint=2 * m, rim_clearance=0.05,
                                 min_separation=0.05))
    return specs


def build_batch(specs, constants):
    """Build every realisation and concatenate them.  ``constants`` (a dict of solver constants)
    is shared by all realisations."""
    dfns = []
    for sp in specs:
        d = build_network(replace(sp, constants=dict(constants)))
        d.consolidate_dfn()
        dfns.append(d)
    K = dfns[0].discharge_int
    fa, ea, offsets = concatenate([(d.fractures_struc_array_hpc, d.elements_struc_array_hpc) for d in dfns])
    return dfns, fa, ea, offsets, K, dfns[0].constants
