"""MI355X-native accelerated solve for AnDFN's analytic-element DFN model."""

from .dfn import (
    DFN,
    BoundingCircle,
    ConstantHeadLine,
    Fracture,
    ImpermeableCircle,
    ImpermeableLine,
    Intersection,
    Well,
)
from .dtypes import dtype_constants, element_dtype_hpc, fracture_dtype_hpc

__version__ = "0.1.0"
