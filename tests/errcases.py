"""Record sets that exercise the solve's error behaviour (shared by the CPU and GPU tests)."""

import numpy as np

import golden_io


def with_isolated_fracture():
    """The template case plus a fifth fracture that holds only its bounding circle.

    The fracture has no discharge element, so its continuity row and its constant's column in
    the discharge system are zero: the matrix is exactly singular (AnDFN's SuperLU raises
    ``RuntimeError``, hpc_solve.py:121).
    """
    fa, ea, consts, K = golden_io.records(golden_io.load("template"))
    nf, ne = len(fa), len(ea)
    fa2 = np.zeros(nf + 1, dtype=fa.dtype)
    fa2[:nf] = fa
    fa2[nf] = fa[0]
    fa2[nf]["_id"] = nf
    fa2[nf]["center"] = fa[0]["center"] + 10.0
    fa2[nf]["elements"] = 0
    fa2[nf]["elements"][0] = ne
    fa2[nf]["nelements"] = 1
    ea2 = np.zeros(ne + 1, dtype=ea.dtype)
    ea2[:ne] = ea
    assert ea[0]["_type"] == 1 and ea[0]["frac0"] == 0
    ea2[ne] = ea[0]
    ea2[ne]["_id"] = ne
    ea2[ne]["frac0"] = nf
    return fa2, ea2, consts, K
