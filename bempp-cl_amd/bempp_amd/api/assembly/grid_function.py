"""Grid functions given by coefficients (a subset of bempp_cl/api/assembly/grid_function.py)."""

import numpy as _np


class GridFunction(object):
    """A function in a space, given by its coefficients (or by projections and a dual space)."""

    def __init__(self, space, coefficients=None, projections=None, dual_space=None):
        self._space = space
        if coefficients is None and projections is None:
            raise ValueError("Either coefficients or projections must be given.")
        if coefficients is not None:
            coefficients = _np.asarray(coefficients)
            if coefficients.shape != (space.global_dof_count,):
                raise ValueError("coefficients must have one entry per global dof")
        self._coefficients = coefficients
        self._projections = None if projections is None else _np.asarray(projections)
        self._dual_space = dual_space

    @property
    def space(self):
        return self._space

    @property
    def coefficients(self):
        if self._coefficients is None:
            raise ValueError("This grid function is given by projections only.")
        return self._coefficients

    @property
    def projections(self):
        return self._projections
