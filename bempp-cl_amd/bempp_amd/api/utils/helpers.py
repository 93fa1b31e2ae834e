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


def get_mass_matrix(domain, dual_to_range):
    """Mass matrix (scipy CSR) of two scalar spaces (bempp_cl/api/utils/helpers.py:88-99)."""
    from bempp_amd.core.sparse_assembler import mass_matrix

    return mass_matrix(domain, dual_to_range)


def get_inverse_mass_matrix(domain, dual_to_range):
    """(Pseudo-)inverse of the mass matrix, as an object with `@` (helpers.py:147-161).

    Square mass matrices are inverted with a sparse LU (scipy SuperLU); a thin one (more rows)
    with the normal equations (M^H M)^{-1} M^H, a thick one with M^H (M M^H)^{-1}, as the
    reference's sparse solver does (discrete_boundary_operator.py:575-646). The factors are
    computed once per pair of spaces and kept.
    """
    key = (domain.id, dual_to_range.id)
    if key not in _INVERSE_MASS:
        _INVERSE_MASS[key] = _InverseSparse(get_mass_matrix(domain, dual_to_range))
    return _INVERSE_MASS[key]


class _InverseSparse(object):
    def __init__(self, mat):
        from scipy.sparse.linalg import splu

        mat = mat.tocsc()
        mh = mat.conjugate().transpose()
        if mat.shape[0] == mat.shape[1]:
            lu = splu(mat)
            self._solve = lu.solve
        elif mat.shape[0] > mat.shape[1]:
            lu = splu((mh @ mat).tocsc())
            self._solve = lambda x: lu.solve(_np.ascontiguousarray(mh @ x))
        else:
            lu = splu((mat @ mh).tocsc())
            self._solve = lambda x: mh @ lu.solve(x)
        self.shape = (mat.shape[1], mat.shape[0])

    def __matmul__(self, x):
        x = _np.asarray(x)
        if _np.iscomplexobj(x):
            return self @ _np.ascontiguousarray(x.real) + 1j * (self @ _np.ascontiguousarray(x.imag))
        return self._solve(_np.ascontiguousarray(x, dtype=_np.float64))


_INVERSE_MASS = {}
