"""Pins of the CPU oracle by what the mathematics fixes (closed forms and invariants).

Each check is chosen so that a plausible slip in the restatement fails it: a wrong branch of
the Joukowsky inverse, a transposed Horner order, a missing halving of the zeroth Fourier
coefficient, a wrong sign in the stream-function fit, a dropped row of the discharge system,
or an iteration that does not reach the fixed point of the discretised problem.
"""

import ctypes

import numpy as np
import pytest

import golden_io
from oracle import oracle

cd_p = ctypes.c_void_p


def _lib():
    lib = oracle.load()
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.oracle_t_map_z_line_to_chi.argtypes = [i64, vp, vp, vp]
    lib.oracle_t_map_chi_to_z_line.argtypes = [i64, vp, vp, vp]
    lib.oracle_t_asym.argtypes = [i64, vp, vp, i64, vp]
    lib.oracle_t_taylor.argtypes = [i64, vp, vp, i64, vp]
    lib.oracle_t_fit.argtypes = [i64, i64, i64, vp, i64, vp]
    lib.oracle_t_calc_error.argtypes = [vp, vp, i64]
    lib.oracle_t_calc_error.restype = ctypes.c_double
    lib.oracle_t_calc_coef_error.argtypes = [vp, vp, i64]
    lib.oracle_t_calc_coef_error.restype = ctypes.c_double
    return lib


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def map_z(z, ep):
    out = np.zeros_like(z)
    _lib().oracle_t_map_z_line_to_chi(len(z), P(z), P(ep), P(out))
    return out


def map_chi(chi, ep):
    out = np.zeros_like(chi)
    _lib().oracle_t_map_chi_to_z_line(len(chi), P(chi), P(ep), P(out))
    return out


EP = np.array([-1.3 + 0.4j, 2.1 - 0.7j])


def test_joukowsky_round_trip_and_exterior_branch():
    rng = np.random.default_rng(1)
    z = (rng.normal(size=400) + 1j * rng.normal(size=400)) * 3
    chi = map_z(z, EP)
    assert np.all(np.abs(chi) >= 1 - 1e-12)  # exterior branch
    assert np.allclose(map_chi(chi, EP), z, rtol=0, atol=1e-12)


def test_joukowsky_segment_maps_to_unit_circle():
    t = np.linspace(0.01, 0.99, 50)
    z = EP[0] + t * (EP[1] - EP[0]) + 1e-13j * (EP[1] - EP[0])
    assert np.allclose(np.abs(map_z(z, EP)), 1.0, atol=1e-6)


def test_joukowsky_far_field():
    z = np.array([1e6 + 3e5j])
    big_z = (2 * z - EP[0] - EP[1]) / (EP[1] - EP[0])
    assert np.allclose(map_z(z, EP), 2 * big_z, rtol=1e-10)


@pytest.mark.parametrize("m", [1, 2, 7, 30])
def test_horner_series_match_explicit_sums(m):
    rng = np.random.default_rng(m)
    c = rng.normal(size=m) + 1j * rng.normal(size=m)
    chi = 1.5 * np.exp(1j * rng.uniform(0, 2 * np.pi, 20))
    out = np.zeros_like(chi)
    _lib().oracle_t_asym(len(chi), P(chi), P(c), m, P(out))
    ref = np.array([sum(complex(c[j]) * complex(x) ** (-j) for j in range(m)) for x in chi])
    assert np.allclose(out, ref, rtol=1e-13)
    _lib().oracle_t_taylor(len(chi), P(chi / 2), P(c), m, P(out))
    ref = np.array([sum(complex(c[j]) * complex(x / 2) ** j for j in range(m)) for x in chi])
    assert np.allclose(out, ref, rtol=1e-13)


@pytest.mark.parametrize("n,m", [(20, 10), (40, 20), (30, 15)])
def test_cosine_fit_is_exact_at_midpoint_nodes(n, m):
    # phi(theta) = sum_j a_j cos(j theta) at theta_i = (i + 1/2) pi / n is reproduced exactly
    rng = np.random.default_rng(n)
    a = rng.normal(size=m)
    th = (np.arange(n) + 0.5) * np.pi / n
    phi = np.array([np.sum(a * np.cos(np.arange(m) * t)) for t in th])
    out = np.zeros(m, dtype=complex)
    _lib().oracle_t_fit(n, m, 0, P(phi), 0, P(out))
    assert np.allclose(out.real, a, atol=1e-13)


@pytest.mark.parametrize("n,m", [(20, 10), (40, 20)])
def test_stream_function_fit_recovers_circle_coefficients(n, m):
    # psi(theta) = Im sum_j c_j e^{ij theta} on theta_i = 2 pi i / n: the fit returns c_j (j >= 1)
    rng = np.random.default_rng(n + 1)
    c = rng.normal(size=m) + 1j * rng.normal(size=m)
    th = 2 * np.pi * np.arange(n) / n
    psi = np.array([np.imag(np.sum(c * np.exp(1j * np.arange(m) * t))) for t in th])
    out = np.zeros(m, dtype=complex)
    _lib().oracle_t_fit(n, m, 1, P(psi), 1, P(out))
    assert np.allclose(out[1:], c[1:], atol=1e-12)
    assert np.isclose(out[0], 1j * c[0].imag, atol=1e-12)


def test_error_measures():
    lib = _lib()
    c = np.array([1, 0.5, 0.25, 0.1, 0.05, 0.01, 0.001, 1e-4, 1e-5, 1e-6], dtype=complex)
    old = c * 0.9
    # tail fraction: | |sum c - c[9:]| / |sum c| - 1 |
    s = np.sum(c)
    assert np.isclose(lib.oracle_t_calc_error(P(c), P(old), 10), abs(abs(s - c[9]) / abs(s + 1e-16) - 1), rtol=1e-14)
    assert np.isclose(lib.oracle_t_calc_coef_error(P(c), P(old), 10), np.mean(np.abs(c - old)), rtol=1e-14)
    tiny = np.zeros(10, dtype=complex)
    assert lib.oracle_t_calc_error(P(c), P(tiny), 10) == 0.0


def _discharge_checks(name):
    g = golden_io.load(name)
    fa, ea, consts, K = golden_io.records(g)
    O = oracle.OracleHandle(fa, ea, K)
    A = O.discharge_matrix()
    b = O.head_rhs()
    import scipy.sparse.linalg as spla

    y = spla.splu(A).solve(b)
    O.close()
    return fa, ea, A, b, y


@pytest.mark.parametrize("case", [c for c in ("template", "template_t", "dfn40") if c in golden_io.CASES])
def test_discharge_system_invariants(case):
    fa, ea, A, b, y = _discharge_checks(case)
    assert np.max(np.abs(A @ y - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))
    de = np.nonzero(np.isin(ea["_type"], (0, 2, 3)))[0]
    pos = {e: j for j, e in enumerate(de)}
    # every fracture's net discharge vanishes (continuity rows)
    for f in range(len(fa)):
        tot = 0.0
        for e in fa["elements"][f, : fa["nelements"][f]]:
            if e in pos:
                s = -1.0 if (ea["_type"][e] == 0 and ea["frac1"][e] == f) else 1.0
                tot += s * y[pos[e]]
        assert abs(tot) <= 1e-10 * np.max(np.abs(y[: len(de)]))


def _fixed_point_map(fa, ea, K, mask_el, lu, O):
    """c -> coefficients after one undamped Jacobi step with the discharges of c."""
    lib = O.lib

    def F(cvec):
        ea["coef"][mask_el] = 0
        ea["coef"][mask_el, : cvec.shape[1]] = cvec
        b = O.head_rhs()
        y = lu.solve(b)
        O.set_discharges(y)
        lib.oracle_element_solver(O.h, 1.0)
        return ea["coef"][mask_el, : cvec.shape[1]].copy()

    return F


@pytest.mark.parametrize("case", ["template_t"])
def test_iteration_reaches_dense_fixed_point(case):
    """At fixed coefficient counts the sweep is affine, c' = M c + g; its fixed point, solved
    densely from M assembled column by column, must equal the iterated solution."""
    import scipy.sparse.linalg as spla

    g = golden_io.load(case)
    fa, ea, consts, K = golden_io.records(g)
    m = 10
    ea["ncoef"][ea["_type"] != 2] = m
    ea["nint"][ea["_type"] != 2] = 2 * m
    fi, ei = fa.copy(), ea.copy()
    c2 = np.array(consts)
    c2["MAX_NCOEF"] = m  # no growth
    c2["MAX_ITERATIONS"] = 400
    c2["MAX_ERROR"] = 1e-15
    info = {}
    oracle.solve(fi, ei, K, c2, info=info)

    mask = ea["_type"] != 2
    O = oracle.OracleHandle(fa, ea, K)
    O.lib.oracle_init(O.h)
    lu = spla.splu(O.discharge_matrix())
    F = _fixed_point_map(fa, ea, K, mask, lu, O)
    nel = int(mask.sum())
    zero = np.zeros((nel, m), dtype=complex)
    g0 = F(zero)
    cols = []
    for k in range(nel * m * 2):
        u = np.zeros(nel * m * 2)
        u[k] = 1.0
        cv = (u[: nel * m] + 1j * u[nel * m:]).reshape(nel, m)
        d = F(cv) - g0
        cols.append(np.concatenate([d.real.ravel(), d.imag.ravel()]))
    M = np.array(cols).T
    rhs = np.concatenate([g0.real.ravel(), g0.imag.ravel()])
    x = np.linalg.solve(np.eye(len(rhs)) - M, rhs)
    cstar = (x[: nel * m] + 1j * x[nel * m:]).reshape(nel, m)
    O.close()
    it = ei["coef"][mask, :m]
    assert np.max(np.abs(it - cstar)) <= 1e-9 * np.max(np.abs(cstar)), (info["nit"], np.max(np.abs(it - cstar)))
