"""ctypes binding of libandfn_mi.so (the C ABI declared in include/andfn_mi.h).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C andfn_amd/csrc``).
There is no fallback: when the library is missing or no GPU is visible, the calls raise.
"""

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libandfn_mi.so"

OK = 0
NOT_CONVERGED = 1
EINVAL = -1
ESINGULAR = -2
ECAPACITY = -3
EHIP = -4
ENODEV = -5

FP64 = 0
MIXED = 1

ADJ_GMRES = 0
ADJ_FIXED_POINT = 1

LU_TRSV = 0
LU_GETRS = 1
LU_INVERSE = 2

FACTORS = {"sparse": 0, "rocsolver": 1, "tournament": 2}

W_JIT = 0  # hpc_fracture.calc_w
W_OBJECT = 1  # Fracture.calc_w (streamline tracking)


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("lu_mode", ctypes.c_int32),
        ("far_field", ctypes.c_double),
        ("growth_schedule", ctypes.c_void_p),
        ("schedule_iterations", ctypes.c_int64),
        ("growth_record", ctypes.c_void_p),
        ("record_iterations", ctypes.c_int64),
        ("head_tol", ctypes.c_double),
        ("factor", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("iteration_log", ctypes.c_void_p),
        ("log_iterations", ctypes.c_int64),
    ]


LOG_FIELDS = 6  # ANDFN_MI_LOG_FIELDS


class Info(ctypes.Structure):
    _fields_ = [
        ("nit", ctypes.c_int64),
        ("n_unknowns", ctypes.c_int64),
        ("n_grown", ctypes.c_int64),
        ("max_error_element", ctypes.c_int64),
        ("error", ctypes.c_double),
        ("error_coef", ctypes.c_double),
        ("error_q", ctypes.c_double),
        ("t_setup", ctypes.c_double),
        ("t_factor", ctypes.c_double),
        ("t_iterate", ctypes.c_double),
        ("t_total", ctypes.c_double),
        ("k_sweep_ms", ctypes.c_double),
        ("k_rhs_ms", ctypes.c_double),
        ("k_fit_ms", ctypes.c_double),
        ("k_lu_ms", ctypes.c_double),
        ("n_sweep", ctypes.c_int64),
        ("n_rhs", ctypes.c_int64),
        ("pairs_sweep", ctypes.c_int64),
        ("pairs_rhs", ctypes.c_int64),
        ("flops_sweep", ctypes.c_double),
        ("flops_rhs", ctypes.c_double),
        ("head_residual", ctypes.c_double),
        ("head_residual_row", ctypes.c_int64),
        ("bc_residual", ctypes.c_double),
        ("nit_fp32", ctypes.c_int64),
        ("solves_fp32", ctypes.c_int64),
        ("refine_rate", ctypes.c_double),
        ("sweep_points_per_lane", ctypes.c_int32),
        ("rhs_points_per_lane", ctypes.c_int32),
        ("nnz_discharge", ctypes.c_int64),
        ("nnz_factors", ctypes.c_int64),
        ("factor_used", ctypes.c_int32),
        ("factor_rounds", ctypes.c_int32),
        ("solve_passes", ctypes.c_int64),
        ("solve_values", ctypes.c_int64),
        ("runahead_iterations", ctypes.c_int64),
        ("graph_iterations", ctypes.c_int64),
        ("sweep_persistent", ctypes.c_int32),
        ("rhs_persistent", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Objective(ctypes.Structure):
    _fields_ = [
        ("dq", ctypes.c_void_p),
        ("dconstant", ctypes.c_void_p),
        ("n_obs", ctypes.c_int64),
        ("obs_frac", ctypes.c_void_p),
        ("obs_z", ctypes.c_void_p),
        ("dhead", ctypes.c_void_p),
    ]


class AdjointOptions(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double),
        ("max_iterations", ctypes.c_int64),
        ("damping", ctypes.c_double),
        ("verbose", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("accept_unconverged", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class AdjointInfo(ctypes.Structure):
    _fields_ = [
        ("nit", ctypes.c_int64),
        ("delta", ctypes.c_double),
        ("objective", ctypes.c_double),
        ("t_setup", ctypes.c_double),
        ("t_iterate", ctypes.c_double),
        ("t_total", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Disks(ctypes.Structure):
    """andfn_mi_disks: disks as structure of arrays ([n][3] centre, normal, x and y vectors; [n] radius)."""
    _fields_ = [
        ("n", ctypes.c_int64),
        ("center", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("x_vector", ctypes.c_void_p),
        ("y_vector", ctypes.c_void_p),
        ("radius", ctypes.c_void_p),
    ]


# every exported symbol and its signature
SIGNATURES = {
    "andfn_mi_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(Options), ctypes.POINTER(Info)]),
    "andfn_mi_heads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int32]),
    "andfn_mi_omega": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int32]),
    "andfn_mi_w": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32]),
    "andfn_mi_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "andfn_mi_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Options), ctypes.POINTER(Info)]),
    "andfn_mi_download": (ctypes.c_int, [ctypes.c_void_p]),
    "andfn_mi_destroy": (None, [ctypes.c_void_p]),
    "andfn_mi_adjoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Objective), ctypes.POINTER(AdjointOptions),
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(AdjointInfo)]),
    "andfn_mi_set_parameters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "andfn_mi_state_heads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "andfn_mi_state_omega": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "andfn_mi_state_w": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]),
    "andfn_mi_state_streamlines": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                                                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "andfn_mi_state_streamlines_fetch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_void_p]),
    "andfn_mi_state_heads_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
    "andfn_mi_state_discharges": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "andfn_mi_candidate_pairs": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Disks), ctypes.c_double, ctypes.c_int64,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "andfn_mi_disk_chords": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(Disks), ctypes.POINTER(Disks), ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "andfn_mi_components": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "andfn_mi_device_count": (ctypes.c_int, []),
    "andfn_mi_last_error": (ctypes.c_char_p, []),
}

_lib = None


class AndfnMiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"andfn_mi error {code}: {msg}")
        self.code = code


def load():
    """Load the in-tree library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C andfn_amd/csrc")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CapacityError(AndfnMiError, IndexError):
    """A record capacity was exceeded (AnDFN raises IndexError there: andfn/dfn.py:790-791)."""


def check(rc):
    if rc == ECAPACITY:
        raise CapacityError(rc, load().andfn_mi_last_error().decode())
    if rc < 0:
        raise AndfnMiError(rc, load().andfn_mi_last_error().decode())
    return rc


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def capacities(fracture_struc_array, element_struc_array):
    return (int(fracture_struc_array.dtype["elements"].shape[0]), int(element_struc_array.dtype["coef"].shape[0]))


def require_records(fracture_struc_array, element_struc_array):
    for a in (fracture_struc_array, element_struc_array):
        if not isinstance(a, np.ndarray) or a.dtype.names is None or not a.flags.c_contiguous:
            raise TypeError("record arrays must be C-contiguous numpy structured arrays")
