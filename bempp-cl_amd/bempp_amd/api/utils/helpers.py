"""Small helpers (bempp_cl/api/utils/helpers.py:49-85)."""

import collections as _collections

import numpy as _np

TypeContainer = _collections.namedtuple("TypeContainer", "real complex opencl")


def get_type(precision):
    """Return (real, complex) numpy type names for 'single' or 'double'."""
    if precision == "single":
        return TypeContainer("float32", "complex64", "float")
    if precision == "double":
        return TypeContainer("float64", "complex128", "double")
    raise ValueError("precision must be one of 'single' or 'double'")


def promote_to_double_precision(array):
    """Convert an array to real or complex double precision."""
    if array.dtype == _np.float32:
        return array.astype(_np.float64, copy=False)
    if array.dtype == _np.complex64:
        return array.astype(_np.complex128, copy=False)
    return array
