"""Load the golden fixtures written by tests/golden/make_golden.py into solver records."""

from pathlib import Path

import numpy as np

from andfn_amd.dtypes import make_element_dtype, make_fracture_dtype

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))


def records(g, max_ncoef=200, max_elements=500):
    """(fractures, elements, constants, discharge_int) as AnDFN consolidated them."""
    ne = len(g["in_e__type"])
    nf = len(g["in_f_t"])
    ea = np.zeros(ne, dtype=make_element_dtype(max_ncoef))
    ea["_id"] = np.arange(ne)
    for k in ("_type", "frac0", "frac1", "endpoints0", "endpoints1", "radius", "center", "head", "phi",
              "ncoef", "nint", "q", "error"):
        ea[k] = g["in_e_" + k]
    nc = g["in_e_coef"].shape[1]
    ea["coef"][:, :nc] = g["in_e_coef"]
    fa = np.zeros(nf, dtype=make_fracture_dtype(max_elements))
    fa["_id"] = np.arange(nf)
    for k in ("t", "radius", "center", "normal", "x_vector", "y_vector", "nelements", "constant"):
        fa[k] = g["in_f_" + k]
    me = g["in_f_elements"].shape[1]
    fa["elements"][:, :me] = g["in_f_elements"]
    constants = np.array(g["constants"])
    return fa, ea, constants, int(g["discharge_int"])


def compare(g, fa, ea):
    """Max relative differences of the converged outputs against a golden run."""
    nc = g["out_coef"].shape[1]
    out = {}
    out["ncoef_equal"] = bool(np.array_equal(ea["ncoef"], g["out_ncoef"]))
    out["nint_equal"] = bool(np.array_equal(ea["nint"], g["out_nint"]))
    c = ea["coef"][:, : max(nc, int(ea["ncoef"].max()))]
    gc = np.zeros_like(c)
    gc[:, :nc] = g["out_coef"]
    scale = np.max(np.abs(gc))
    out["coef"] = float(np.max(np.abs(c - gc)) / scale)
    out["q"] = float(np.max(np.abs(ea["q"] - g["out_q"])) / np.max(np.abs(g["out_q"])))
    out["constant"] = float(np.max(np.abs(fa["constant"] - g["out_constant"])) / np.max(np.abs(g["out_constant"])))
    return out


def compare_errors(g, ea):
    """The in-place error outputs of a solve against a golden run (hpc_math_functions.py:611-670,
    406-447):

    * ``error_coef`` (mean |c_j - c_old_j| per element) and ``dpsi_corr`` (branch-cut corrections of
      the bounding circles): max absolute difference relative to the largest golden coefficient and
      the largest golden correction; both are differences or sums of coefficient-sized terms, so
      their rounding floor is set by those magnitudes;
    * ``error`` = | |sum c[:0.9m]| / |sum c| - 1 | per element: the largest ratio of the difference
      to its first-order conditioning bound m * delta * (|S| + |H|) / |S|^2 at a coefficient
      agreement delta = 1e-9 of the largest golden coefficient (S = sum c, H = sum c[:0.9m] of the
      golden coefficients); below 1 means the difference is what 1e-9 coefficient agreement allows.
    """
    gc = g["out_coef"]
    m = g["out_ncoef"]
    scale = float(np.max(np.abs(gc)))
    out = {}
    out["error_coef"] = float(np.max(np.abs(ea["error_coef"] - g["out_error_coef"])) / scale)
    gd = g["out_dpsi_corr"]
    dscale = float(np.max(np.abs(gd)))
    out["dpsi_corr"] = float(np.max(np.abs(ea["dpsi_corr"][:, : gd.shape[1]] - gd)) / (dscale if dscale > 0 else 1.0))
    worst = 0.0
    ge = g["out_error"]
    for i in range(len(ge)):
        mi = int(m[i])
        c = gc[i, :mi]
        S = c.sum()
        H = c[: int(0.9 * mi)].sum()
        if mi < 1 or abs(S) == 0.0:
            continue
        bound = 1e-9 * scale * mi * (abs(S) + abs(H)) / abs(S) ** 2
        worst = max(worst, abs(float(ea["error"][i]) - float(ge[i])) / bound)
    out["error_ratio"] = worst
    return out
