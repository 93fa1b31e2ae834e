"""Potential operators (bempp_cl/api/assembly/potential_operator.py:4-91)."""


class PotentialOperator(object):
    """Evaluates a potential of a GridFunction at fixed points."""

    def __init__(self, assembler):
        self._assembler = assembler

    def evaluate(self, grid_fun):
        """Return the potential values, shape (kernel_dimension, npoints)."""
        if not grid_fun.space.is_compatible(self._assembler.space):
            raise ValueError("GridFunction space and potential space are not compatible.")
        return self._assembler.evaluate(grid_fun.coefficients)

    def __mul__(self, obj):
        from bempp_amd.api.assembly.grid_function import GridFunction

        if isinstance(obj, GridFunction):
            return self.evaluate(obj)
        return NotImplemented

    @property
    def space(self):
        return self._assembler.space

    @property
    def kernel_dimension(self):
        return self._assembler.kernel_dimension

    @property
    def descriptor(self):
        return self._assembler.descriptor
