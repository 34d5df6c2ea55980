"""Record concatenation of independent realisations (CPU)."""

import numpy as np

from andfn_amd.batch import concatenate, realisation_specs, split
from andfn_amd.generator import build_network, workload


def _records(name, **kw):
    d = build_network(workload(name, **kw))
    d.consolidate_dfn()
    return d.fractures_struc_array_hpc, d.elements_struc_array_hpc


def test_concatenate_offsets_links_and_split_round_trips():
    nets = [_records("dfn40"), _records("dfn40w"), _records("dfn40", seed=5)]
    fa, ea, off = concatenate(nets)
    assert off[-1] == (sum(len(f) for f, _ in nets), sum(len(e) for _, e in nets))
    assert np.array_equal(ea["_id"], np.arange(len(ea)))
    assert np.array_equal(fa["_id"], np.arange(len(fa)))
    for r, (fb, eb) in enumerate(split(fa, ea, off)):
        f0, e0 = off[r]
        fr, er = nets[r]
        assert np.array_equal(eb["frac0"], er["frac0"] + f0)
        inter = er["_type"] == 0
        assert np.array_equal(eb["frac1"][inter], er["frac1"][inter] + f0)
        assert np.array_equal(eb["frac1"][~inter], er["frac1"][~inter])
        for f in range(len(fr)):
            n = fr["nelements"][f]
            assert np.array_equal(fb["elements"][f, :n], fr["elements"][f, :n] + e0)
        for k in ("coef", "endpoints0", "endpoints1", "phi", "ncoef", "nint"):
            assert np.array_equal(eb[k], er[k])
        assert np.array_equal(fb["t"], fr["t"])


def test_realisation_family_is_seeded_and_in_range():
    a = realisation_specs(16)
    b = realisation_specs(16)
    assert [(s.n_fractures, s.ncoef, s.seed) for s in a] == [(s.n_fractures, s.ncoef, s.seed) for s in b]
    for s in a:
        assert 500 <= s.n_fractures <= 4000 and 10 <= s.ncoef <= 20 and s.nint == 2 * s.ncoef
        assert abs(s.cube - 100.0 * (s.n_fractures / 200.0) ** (1.0 / 3.0)) < 1e-9
