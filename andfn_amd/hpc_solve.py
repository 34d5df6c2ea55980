"""Drop-in replacement of AnDFN's HPC solve entry point.

``solve(fracture_struc_array, element_struc_array, discharge_int, constants, ntype_elements)``
has the signature and in-place semantics of ``andfn.hpc.hpc_solve.solve``
(``andfn/hpc/hpc_solve.py:55-257``): the element records (coef, ncoef, nint, thetas, q,
dpsi_corr and the error fields) and the fracture constants are updated in place and the
element array is returned.  In an environment that has AnDFN, assigning
``andfn.dfn.hpc_solve = andfn_amd.hpc_solve.solve`` routes ``DFN.solve`` here.

The solve runs on the GPU through ``libandfn_mi.so``; there is no CPU path.  Like the
reference, a solve that stops at MAX_ITERATIONS returns normally (a warning is logged); a
singular discharge matrix raises ``RuntimeError`` (SuperLU's behaviour in the reference).
"""

import ctypes
import logging

import numpy as np

from . import _lib

logger = logging.getLogger("andfn")
_last_info = None


def constant(constants, name):
    """Scalar field ``name`` of a dtype_constants record (0-d, or a one-element array)."""
    return np.asarray(constants)[name].reshape(-1)[0].item()


def _log_solve(nf, info, log_rows, error_q_final):
    """AnDFN's log lines of one solve on the ``andfn`` logger (hpc_solve.py:140-141, 198-216, 242-255)."""
    logger.info(f"Number of fractures: {nf}")
    logger.info(f"Number of entries in discharge matrix: {info.nnz_discharge}")
    for k, row in enumerate(log_rows):
        nit = k + 1
        error, error_coef, error_q, el, nc, ty = row
        it = f"0{nit}" if nit < 10 else f"{nit}"
        logger.info(f"Iteration: {it}, Error E: {error:.3e}, Coef: {error_coef:.3e}, Q: {error_q:.3e}, "
                    f"Element: {int(el)}, N of Coef: {int(nc)}, Type: {int(ty)}")
    logger.info("---------------------------------------")
    logger.info("Solver results")
    logger.info("---------------------------------------")
    logger.info(f"Iterations: {info.nit}, Error E: {info.error:.3e}, Q: {error_q_final:.3e}, ")
    days, rem = divmod(info.t_total, 86400)
    hours, rem = divmod(rem, 3600)
    minutes, seconds = divmod(rem, 60)
    logger.info(f"Solve time: {int(days)} days, {int(hours)} hours, {int(minutes)} minutes, {seconds:.2f} seconds\n")


def last_solve_info():
    """Dict of the last solve's iteration count, errors, timings and kernel counters."""
    return dict(_last_info) if _last_info is not None else None


def solve(fracture_struc_array, element_struc_array, discharge_int, constants, ntype_elements=None, precision="fp64",
          device=0, verbose=False, lu_mode=0, growth_schedule=None, record_growth=False, head_tol=0.0,
          factor="sparse"):
    """Solve in place; see the module docstring.

    ``growth_schedule`` (int32 [iterations, n_elements], optional) replays another run's
    coefficient counts; ``record_growth`` stores this run's counts per iteration in
    ``last_solve_info()["ncoef_history"]``.  ``head_tol`` > 0 replaces AnDFN's stop rule: the
    solve stops as soon as the head residual (max |A y - b(c)| over the head equations, metres)
    is at most ``head_tol`` (or at MAX_ITERATIONS).  ``factor``: "sparse" (the default: the
    library's multifrontal LU of the sparse discharge matrix, as the reference factors it with
    SuperLU), or the dense "tournament" (blocked LU with tournament pivoting, all connected parts
    of the network in one batch) or "rocsolver" (rocSOLVER getrf).
    """
    global _last_info
    _lib.require_records(fracture_struc_array, element_struc_array)
    lib = _lib.load()
    me, m = _lib.capacities(fracture_struc_array, element_struc_array)
    consts = np.ascontiguousarray(constants)
    if consts.dtype.names is None or "MAX_ERROR" not in consts.dtype.names:
        raise TypeError("constants must be a dtype_constants record")
    nte = np.zeros(6, dtype=np.int64) if ntype_elements is None else np.ascontiguousarray(ntype_elements, dtype=np.int64)
    opt = _options(precision, device, verbose, lu_mode, head_tol, factor)
    sched = None
    if growth_schedule is not None:
        sched = np.ascontiguousarray(growth_schedule, dtype=np.int32)
        if sched.ndim != 2 or sched.shape[1] != len(element_struc_array):
            raise ValueError("growth_schedule must have shape (iterations, n_elements)")
        opt.growth_schedule = sched.ctypes.data
        opt.schedule_iterations = sched.shape[0]
    max_iterations = int(constant(consts, "MAX_ITERATIONS"))
    rec = None
    if record_growth:
        rec = np.zeros((max(max_iterations, 0), len(element_struc_array)), dtype=np.int32)
        opt.growth_record = rec.ctypes.data
        opt.record_iterations = rec.shape[0]
    log = None
    if logger.isEnabledFor(logging.INFO):
        log = np.zeros((max(max_iterations, 0), _lib.LOG_FIELDS), dtype=np.float64)
        opt.iteration_log = log.ctypes.data
        opt.log_iterations = log.shape[0]
    info = _lib.Info()
    rc = lib.andfn_mi_solve(_lib.ptr(fracture_struc_array), len(fracture_struc_array), me,
                            _lib.ptr(element_struc_array), len(element_struc_array), m, int(discharge_int),
                            _lib.ptr(consts), _lib.ptr(nte), ctypes.byref(opt), ctypes.byref(info))
    if rc == _lib.ESINGULAR:
        raise RuntimeError("Factor is exactly singular: " + lib.andfn_mi_last_error().decode())
    _lib.check(rc)
    _last_info = info.as_dict()
    _last_info["converged"] = rc == _lib.OK
    if rec is not None:
        _last_info["ncoef_history"] = rec[: info.nit].copy()
    if log is not None:
        # AnDFN sets error_q to 1e30 on each stop-rule hit (hpc_solve.py:237-239), so a run ended
        # by that rule logs 1e30
        eq = 1e30 if (rc == _lib.OK and head_tol <= 0.0) else info.error_q
        _log_solve(len(fracture_struc_array), info, log[: info.nit], eq)
    if rc == _lib.NOT_CONVERGED:
        logger.warning(f"solve stopped at MAX_ITERATIONS ({info.nit}) without meeting the tolerance")
    return element_struc_array


def get_heads(fracture_struc_array, element_struc_array, frac, z, device=0):
    """Heads h = Re Omega_f(z)/t_f at points z (local complex coordinates) on fractures ``frac``
    (``andfn/hpc/hpc_fracture.py:204-250``)."""
    _lib.require_records(fracture_struc_array, element_struc_array)
    lib = _lib.load()
    me, m = _lib.capacities(fracture_struc_array, element_struc_array)
    frac = np.ascontiguousarray(frac, dtype=np.int64)
    zz = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(len(zz), dtype=np.float64)
    _lib.check(lib.andfn_mi_heads(_lib.ptr(fracture_struc_array), len(fracture_struc_array), me,
                                  _lib.ptr(element_struc_array), len(element_struc_array), m, len(zz), _lib.ptr(frac),
                                  _lib.ptr(zz), _lib.ptr(out), int(device)))
    return out


def get_omega(fracture_struc_array, element_struc_array, frac, z, device=0):
    """Complex potential Omega_f(z) = C_f + the fracture's elements at points z (local complex
    coordinates) on fractures ``frac`` (``andfn/hpc/hpc_fracture.py:51-92``, calc_omega)."""
    _lib.require_records(fracture_struc_array, element_struc_array)
    lib = _lib.load()
    me, m = _lib.capacities(fracture_struc_array, element_struc_array)
    frac = np.ascontiguousarray(frac, dtype=np.int64)
    zz = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(len(zz), dtype=np.complex128)
    _lib.check(lib.andfn_mi_omega(_lib.ptr(fracture_struc_array), len(fracture_struc_array), me,
                                  _lib.ptr(element_struc_array), len(element_struc_array), m, len(zz), _lib.ptr(frac),
                                  _lib.ptr(zz), _lib.ptr(out), int(device)))
    return out


def _w_form(form):
    forms = {"jit": _lib.W_JIT, "object": _lib.W_OBJECT}
    if form not in forms:
        raise ValueError(f"form must be one of {sorted(forms)}")
    return forms[form]


def get_w(fracture_struc_array, element_struc_array, frac, z, exclude=-1, form="jit", device=0):
    """Complex discharge vector w = -dOmega/dz at points z (local complex coordinates) on fractures
    ``frac``, every element but ``exclude`` summed.  ``form="jit"`` is ``hpc_fracture.calc_w``
    (``andfn/hpc/hpc_fracture.py:139-174``: no impermeable objects, the bounding circle as
    ``taylor_series_d1`` alone); ``form="object"`` is ``Fracture.calc_w`` (``andfn/fracture.py:453-477``),
    the form the streamline tracking uses."""
    _lib.require_records(fracture_struc_array, element_struc_array)
    lib = _lib.load()
    me, m = _lib.capacities(fracture_struc_array, element_struc_array)
    frac = np.ascontiguousarray(frac, dtype=np.int64)
    zz = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(len(zz), dtype=np.complex128)
    _lib.check(lib.andfn_mi_w(_lib.ptr(fracture_struc_array), len(fracture_struc_array), me,
                              _lib.ptr(element_struc_array), len(element_struc_array), m, len(zz), _lib.ptr(frac),
                              _lib.ptr(zz), int(exclude), _w_form(form), _lib.ptr(out), int(device)))
    return out


def _options(precision="fp64", device=0, verbose=False, lu_mode=0, head_tol=0.0, factor="sparse"):
    if factor not in _lib.FACTORS:
        raise ValueError(f"factor must be one of {sorted(_lib.FACTORS)}")
    return _lib.Options(_lib.MIXED if precision == "mixed" else _lib.FP64, int(device), 1 if verbose else 0,
                        int(lu_mode), 0.0, None, 0, None, 0, float(head_tol), _lib.FACTORS[factor], 0, None, 0)


class DeviceNetwork:
    """A network uploaded once for repeated solves (C ABI create/run/download/destroy).

    Every ``run`` restarts from the uploaded initial state with the inputs already resident in
    device memory; ``download`` writes the last run's results into the record arrays.
    """

    def __init__(self, fracture_struc_array, element_struc_array, discharge_int, constants, device=0):
        self.h = None
        _lib.require_records(fracture_struc_array, element_struc_array)
        self.lib = _lib.load()
        self.fa = fracture_struc_array
        self.ea = element_struc_array
        self.consts = np.ascontiguousarray(constants)
        me, m = _lib.capacities(fracture_struc_array, element_struc_array)
        self.h = ctypes.c_void_p()
        _lib.check(self.lib.andfn_mi_create(_lib.ptr(self.fa), len(self.fa), me, _lib.ptr(self.ea), len(self.ea), m,
                                            int(discharge_int), _lib.ptr(self.consts), int(device),
                                            ctypes.byref(self.h)))
        self.device = device

    def run(self, precision="fp64", lu_mode=0, head_tol=0.0, verbose=False, factor="sparse"):
        opt = _options(precision, self.device, verbose, lu_mode, head_tol, factor)
        info = _lib.Info()
        rc = self.lib.andfn_mi_run(self.h, ctypes.byref(opt), ctypes.byref(info))
        _lib.check(rc)
        d = info.as_dict()
        d["converged"] = rc == _lib.OK
        return d

    def download(self):
        _lib.check(self.lib.andfn_mi_download(self.h))

    def set_parameters(self, t=None, head=None):
        """Replace fracture transmissivities t[F] and/or prescribed heads head[E] of the
        constant-head elements (phi = head * t); the next ``run`` factors the new matrix."""
        tt = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
        hh = None if head is None else np.ascontiguousarray(head, dtype=np.float64)
        if tt is not None and tt.shape != (len(self.fa),):
            raise ValueError("t must have one value per fracture")
        if hh is not None and hh.shape != (len(self.ea),):
            raise ValueError("head must have one value per element")
        _lib.check(self.lib.andfn_mi_set_parameters(self.h, None if tt is None else _lib.ptr(tt),
                                                    None if hh is None else _lib.ptr(hh)))

    def state_discharges(self):
        """(q[E], constant[F]) of the current state."""
        q = np.zeros(len(self.ea))
        c = np.zeros(len(self.fa))
        _lib.check(self.lib.andfn_mi_state_discharges(self.h, _lib.ptr(q), _lib.ptr(c)))
        return q, c

    def state_heads(self, frac, z):
        """Heads at points (fracture index, local complex coordinate) of the current state."""
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        zz = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros(len(zz))
        _lib.check(self.lib.andfn_mi_state_heads(self.h, len(zz), _lib.ptr(frac), _lib.ptr(zz), _lib.ptr(out)))
        return out

    def state_omega(self, frac, z):
        """Complex potential at points (fracture index, local complex coordinate) of the current state."""
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        zz = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros(len(zz), dtype=np.complex128)
        _lib.check(self.lib.andfn_mi_state_omega(self.h, len(zz), _lib.ptr(frac), _lib.ptr(zz), _lib.ptr(out)))
        return out

    def state_w(self, frac, z, exclude=-1, form="object"):
        """Discharge vector w at points of the current state (``get_w``'s forms)."""
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        zz = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros(len(zz), dtype=np.complex128)
        _lib.check(self.lib.andfn_mi_state_w(self.h, len(zz), _lib.ptr(frac), _lib.ptr(zz), int(exclude),
                                             _w_form(form), _lib.ptr(out)))
        return out

    def streamlines(self, frac, z0, elevation, aperture, ds, max_length, backward=False):
        """Streamline tracking of many lines on the current state (``andfn_mi_state_streamlines``).

        Returns (points, speeds, n_points [n], segment_fracture, segment_points, n_segments [n],
        exit_element [n]): points and speeds of every line back to back, the segments' fractures and
        lengths back to back, exit_element -1 where a line left through no element."""
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        z0 = np.ascontiguousarray(z0, dtype=np.complex128)
        el = np.ascontiguousarray(elevation, dtype=np.float64)
        ap = np.ascontiguousarray(aperture, dtype=np.float64)
        n = len(frac)
        if len(z0) != n or len(el) != n or ap.shape != (len(self.fa),):
            raise ValueError("one start and elevation per line; one aperture per fracture")
        npts = np.zeros(n, dtype=np.int64)
        nseg = np.zeros(n, dtype=np.int64)
        ex = np.zeros(n, dtype=np.int64)
        _lib.check(self.lib.andfn_mi_state_streamlines(self.h, n, _lib.ptr(frac), _lib.ptr(z0), _lib.ptr(el),
                                                       _lib.ptr(ap), float(ds), 1 if backward else 0, int(max_length),
                                                       _lib.ptr(npts), _lib.ptr(nseg), _lib.ptr(ex)))
        pts = np.zeros(int(npts.sum()), dtype=np.complex128)
        vel = np.zeros(len(pts), dtype=np.float64)
        sf = np.zeros(int(nseg.sum()), dtype=np.int64)
        sp = np.zeros(len(sf), dtype=np.int64)
        _lib.check(self.lib.andfn_mi_state_streamlines_fetch(self.h, _lib.ptr(pts), _lib.ptr(vel), _lib.ptr(sf),
                                                             _lib.ptr(sp)))
        return pts, vel, npts, sf, sp, nseg, ex

    def adjoint(self, dq=None, dconstant=None, obs_frac=None, obs_z=None, dhead=None, tol=1e-10,
                max_iterations=5000, damping=0.75, verbose=False, method="gmres", accept_unconverged=False):
        """Gradient of J = dq.q + dconstant.C + dhead.h(obs) at the last run's state with respect
        to every fracture transmissivity and every prescribed head (implicit adjoint of the
        fixed point).  Returns (dJ/dt [F], dJ/dhead [E], info).  A run that stopped at
        MAX_ITERATIONS without meeting its stop rule is refused unless ``accept_unconverged``
        (the adjoint is the gradient of the fixed point, not of the iterate)."""
        keep = []

        def arr(a, dtype, n=None):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            if n is not None and a.shape[0] != n:
                raise ValueError("array length mismatch")
            keep.append(a)
            return a.ctypes.data

        nobs = 0 if obs_frac is None else len(obs_frac)
        obj = _lib.Objective(arr(dq, np.float64, len(self.ea)), arr(dconstant, np.float64, len(self.fa)), nobs,
                             arr(obs_frac, np.int64, nobs) if nobs else None,
                             arr(obs_z, np.complex128, nobs) if nobs else None,
                             arr(dhead, np.float64, nobs) if nobs else None)
        if method not in ("gmres", "fixed_point"):
            raise ValueError("method must be 'gmres' or 'fixed_point'")
        opt = _lib.AdjointOptions(float(tol), int(max_iterations), float(damping), 1 if verbose else 0,
                                  _lib.ADJ_GMRES if method == "gmres" else _lib.ADJ_FIXED_POINT,
                                  1 if accept_unconverged else 0, 0)
        gt = np.zeros(len(self.fa))
        gh = np.zeros(len(self.ea))
        info = _lib.AdjointInfo()
        rc = self.lib.andfn_mi_adjoint(self.h, ctypes.byref(obj), ctypes.byref(opt), _lib.ptr(gt), _lib.ptr(gh),
                                       ctypes.byref(info))
        _lib.check(rc)
        d = info.as_dict()
        d["converged"] = rc == _lib.OK
        return gt, gh, d

    def close(self):
        if self.h:
            self.lib.andfn_mi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
