"""Record layouts shared with AnDFN's HPC path.

The solver exchanges data with its caller through numpy structured arrays whose
byte layout is identical to AnDFN's, so arrays built here and arrays built by
AnDFN are interchangeable:

* ``element_dtype_hpc``  follows ``andfn/element.py:45-69``
* ``fracture_dtype_hpc`` follows ``andfn/element.py:146-159``
* ``dtype_constants``    follows ``andfn/constants.py:13-29``

The fixed capacities ``MAX_NCOEF`` (coefficients per element, default 200) and
``MAX_ELEMENTS`` (elements per fracture, default 500) are the defaults of
``andfn/element.py:13-20``; ``make_element_dtype``/``make_fracture_dtype``
build the layouts for other capacities.
"""

import numpy as np

MAX_NCOEF = 200
MAX_ELEMENTS = 500

# element types (andfn/element.py:231-238)
INTERSECTION = 0
BOUNDING_CIRCLE = 1
WELL = 2
CONST_HEAD_LINE = 3
IMP_CIRCLE = 4
IMP_LINE = 5


def make_element_dtype(max_ncoef=MAX_NCOEF):
    return np.dtype(
        [
            ("_id", np.int64),
            ("_type", np.int64),
            ("frac0", np.int64),
            ("frac1", np.int64),
            ("endpoints0", np.complex128, 2),
            ("endpoints1", np.complex128, 2),
            ("radius", np.float64),
            ("center", np.complex128),
            ("head", np.float64),
            ("phi", np.float64),
            ("ncoef", np.int64),
            ("nint", np.int64),
            ("q", np.float64),
            ("thetas", np.float64, max_ncoef * 2),
            ("coef", np.complex128, max_ncoef),
            ("old_coef", np.complex128, max_ncoef),
            ("dpsi_corr", np.float64, max_ncoef * 2),
            ("error", np.float64),
            ("error_old", np.float64),
            ("error_old2", np.float64),
            ("error_coef", np.float64),
        ]
    )


def make_fracture_dtype(max_elements=MAX_ELEMENTS):
    return np.dtype(
        [
            ("_id", np.int64),
            ("t", np.float64),
            ("radius", np.float64),
            ("center", np.float64, 3),
            ("normal", np.float64, 3),
            ("x_vector", np.float64, 3),
            ("y_vector", np.float64, 3),
            ("elements", np.int64, max_elements),
            ("nelements", np.int64),
            ("constant", np.float64),
        ]
    )


element_dtype_hpc = make_element_dtype()
fracture_dtype_hpc = make_fracture_dtype()

dtype_constants = np.dtype(
    [
        ("RHO", np.float64),
        ("G", np.float64),
        ("SE_FACTOR", np.float64),
        ("MAX_ITERATIONS", np.int32),
        ("MAX_ERROR", np.float64),
        ("MAX_NCOEF", np.int32),
        ("DAMPING", np.float64),
        ("COEF_INCREASE", np.int32),
        ("COEF_RATIO", np.float64),
        ("MAX_ELEMENTS", np.int32),
        ("NCOEF", np.int32),
        ("NINT", np.int32),
        ("NUM_THREADS", np.int32),
    ]
)


def default_constants(**overrides):
    """0-d constants record with AnDFN's defaults (``andfn/constants.py:89-106``)."""
    c = np.array(
        (1000.0, 9.81, 0.9, 50, 1e-6, 150, 0.5, 5, 0.05, 150, 5, 10, -1),
        dtype=dtype_constants,
    )
    for k, v in overrides.items():
        if k not in dtype_constants.names:
            raise KeyError(f"unknown constant {k}")
        c[k] = v
    return c


def element_capacity(element_struc_array):
    """Coefficient capacity (MAX_NCOEF) of an element record array."""
    return int(element_struc_array.dtype["coef"].shape[0])


def fracture_capacity(fracture_struc_array):
    """Element-list capacity (MAX_ELEMENTS) of a fracture record array."""
    return int(fracture_struc_array.dtype["elements"].shape[0])
