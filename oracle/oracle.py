"""Driver of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this module; the
product path never does.  It restates the outer loop of AnDFN's ``hpc_solve.solve``
(``andfn/hpc/hpc_solve.py:55-257``) over the C++ kernels of ``aem_oracle.cpp``:

* discharge matrix assembly, ``scipy.sparse.csc_matrix`` with duplicates summed, and
  ``scipy.sparse.linalg.splu`` factorisation (hpc_solve.py:110-121, 832-887);
* per iteration: head right-hand side, SuperLU solve, scatter of q and the fracture
  constants, relative discharge change (hpc_solve.py:155-174, 594-663);
* the element sweep with damping and coefficient growth (hpc_solve.py:179-191, 393-591);
* the extra sweep at the iteration cap and the two-hit stop rule (hpc_solve.py:224-239).
"""

import ctypes
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = HERE / "liboracle_aem.so"
    if not path.exists():
        raise RuntimeError(f"oracle library not built: run make -C {HERE}")
    lib = ctypes.CDLL(str(path))
    vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [vp, i64, i64, vp, i64, i64, i64]
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_num_discharge.restype = i64
    lib.oracle_num_discharge.argtypes = [vp]
    lib.oracle_discharge_nnz.restype = i64
    lib.oracle_discharge_nnz.argtypes = [vp]
    lib.oracle_discharge_matrix.argtypes = [vp, vp, vp, vp]
    lib.oracle_init.argtypes = [vp]
    lib.oracle_old_discharges.argtypes = [vp, vp]
    lib.oracle_head_rhs.argtypes = [vp, vp]
    lib.oracle_set_discharges.argtypes = [vp, vp]
    lib.oracle_element_solver.argtypes = [vp, dbl]
    lib.oracle_element_solver2.restype = i64
    lib.oracle_element_solver2.argtypes = [vp, dbl, i64, dbl, i64, i64, dbl, vp]
    lib.oracle_ncoef.argtypes = [vp, vp]
    lib.oracle_errors.argtypes = [vp, vp]
    lib.oracle_omega_at.argtypes = [vp, i64, vp, vp, i64, vp]
    lib.oracle_w_at.argtypes = [vp, i64, vp, vp, i64, i64, vp]
    _LIB = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleHandle:
    """Owns the C++ oracle state over a pair of record arrays (mutated in place)."""

    def __init__(self, fracture_struc_array, element_struc_array, discharge_int):
        assert fracture_struc_array.flags.c_contiguous and element_struc_array.flags.c_contiguous
        self.lib = load()
        self.fa = fracture_struc_array
        self.ea = element_struc_array
        m = element_struc_array.dtype["coef"].shape[0]
        me = fracture_struc_array.dtype["elements"].shape[0]
        self.h = self.lib.oracle_create(_ptr(self.fa), len(self.fa), me, _ptr(self.ea), len(self.ea), m,
                                        int(discharge_int))
        self.nde = self.lib.oracle_num_discharge(self.h)
        self.n = self.nde + len(self.fa)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def discharge_matrix(self):
        import scipy.sparse as sps

        nnz = self.lib.oracle_discharge_nnz(self.h)
        rows = np.zeros(nnz, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int64)
        data = np.zeros(nnz, dtype=np.float64)
        self.lib.oracle_discharge_matrix(self.h, _ptr(rows), _ptr(cols), _ptr(data))
        return sps.csc_matrix((data, (rows, cols)), shape=(self.n, self.n))

    def head_rhs(self):
        b = np.zeros(self.n)
        self.lib.oracle_head_rhs(self.h, _ptr(b))
        return b

    def set_discharges(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        self.lib.oracle_set_discharges(self.h, _ptr(y))

    def omega_at(self, frac, z, exclude=-1):
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        z = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros(len(z), dtype=np.complex128)
        self.lib.oracle_omega_at(self.h, len(z), _ptr(frac), _ptr(z), int(exclude), _ptr(out))
        return out

    def w_at(self, frac, z, exclude=-1, jit=False):
        """Complex discharge vector w at points: ``jit`` follows hpc_fracture.calc_w
        (hpc_fracture.py:139-174), otherwise the object model's Fracture.calc_w (fracture.py:453-477)."""
        frac = np.ascontiguousarray(frac, dtype=np.int64)
        z = np.ascontiguousarray(z, dtype=np.complex128)
        out = np.zeros(len(z), dtype=np.complex128)
        self.lib.oracle_w_at(self.h, len(z), _ptr(frac), _ptr(z), int(exclude), int(bool(jit)), _ptr(out))
        return out


def solve(fracture_struc_array, element_struc_array, discharge_int, constants, ntype_elements=None, info=None,
          growth_schedule=None, refine=0):
    """CPU restatement of ``andfn.hpc.hpc_solve.solve``; mutates and returns the element array.

    ``info`` (a dict, optional) receives nit, the per-iteration (error, error_coef, error_q)
    history, every element's ncoef after each iteration, and phase timings.
    ``growth_schedule`` (int32 [iterations, n_elements], optional) replays a recorded
    coefficient-count trajectory instead of applying the growth test.
    ``refine`` > 0 follows every SuperLU solve with that many steps of iterative refinement
    (x += LU^-1 (b - A x), residual in FP64): the discharges of the exact solve, to separate the
    iteration's arithmetic from SuperLU's own solve error (not part of the reference, which uses
    ``lu.solve`` alone: hpc_solve.py:651).
    """
    import scipy.sparse.linalg as spla

    max_error = float(constants["MAX_ERROR"])
    max_iterations = int(constants["MAX_ITERATIONS"])
    coef_ratio = float(constants["COEF_RATIO"])
    max_coef = int(constants["MAX_NCOEF"])
    coef_increase = int(constants["COEF_INCREASE"])
    damping = float(constants["DAMPING"])

    t0 = time.perf_counter()
    O = OracleHandle(fracture_struc_array, element_struc_array, discharge_int)
    lib, h = O.lib, O.h
    y = np.zeros(O.n)
    lib.oracle_old_discharges(h, _ptr(y))
    A = O.discharge_matrix()
    t_asm = time.perf_counter()
    lu = spla.splu(A)
    t_lu = time.perf_counter()
    lib.oracle_init(h)

    nit = 0
    cnt_error = 0
    error_q = 1.0
    hist = []
    errs = np.zeros(3)
    grown_total = 0
    ncoef_hist = []
    nc = np.zeros(len(element_struc_array), dtype=np.int32)
    sched = None if growth_schedule is None else np.ascontiguousarray(growth_schedule, dtype=np.int32)
    while cnt_error < 2 and nit < max_iterations:
        nit += 1
        if error_q > max_error / 1e30:
            y_old = y.copy()
            b = O.head_rhs()
            y = lu.solve(b)
            for _ in range(int(refine)):
                y = y + lu.solve(b - A @ y)
            O.set_discharges(y)
            error_q = float(np.max(np.abs(y - y_old) / (np.max(np.abs(y_old)) + 1e-16)))
        force = None
        if sched is not None and nit - 1 < len(sched):
            force = sched[nit - 1].ctypes.data_as(ctypes.c_void_p)
        grown_total += lib.oracle_element_solver2(h, max_error, nit, coef_ratio, max_coef, coef_increase, damping,
                                                  force)
        lib.oracle_ncoef(h, _ptr(nc))
        ncoef_hist.append(nc.copy())
        lib.oracle_errors(h, _ptr(errs))
        error, arg, error_coef = float(errs[0]), int(errs[1]), float(errs[2])
        hist.append((error, error_coef, error_q, arg))
        if nit == max_iterations:
            lib.oracle_element_solver(h, damping)
        if error_coef < max_error and error_q < max_error:
            cnt_error += 1
            error_q = 1e30
    t_end = time.perf_counter()
    if info is not None:
        # head residual of the returned state (max |A y - b(c)| over the head equations, metres)
        b = O.head_rhs()
        O.set_discharges(y)
        nde = O.nde
        ay = A @ y
        scale = np.ones(nde)
        de = np.nonzero(np.isin(element_struc_array["_type"], (0, 2, 3)))[0]
        chl = element_struc_array["_type"][de] != 0
        scale[chl] = 1.0 / fracture_struc_array["t"][element_struc_array["frac0"][de[chl]]]
        info["head_residual"] = float(np.max(np.abs(ay[:nde] - b[:nde]) * scale)) if nde else 0.0
    O.close()
    if info is not None:
        info.update(nit=nit, history=np.array(hist), grown=grown_total, ncoef_history=np.array(ncoef_hist), t_assemble=t_asm - t0,
                    t_factor=t_lu - t_asm, t_iterate=t_end - t_lu, t_total=t_end - t0, n=O.n, nnz=A.nnz)
    return element_struc_array
