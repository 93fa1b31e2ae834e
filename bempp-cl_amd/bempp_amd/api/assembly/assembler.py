"""Assembler selection (bempp_cl/api/assembly/assembler.py:4-186).

"dense" and "default_nonlocal" select the dense assembler, as in the reference
(assembler.py:24-27), and "sparse" the host assembler of the L2 identity; other assembler
types of the reference (fmm, singular-only, diagonal) are outside this library and raise
ValueError.
"""


def _create_assembler(domain, dual_to_range, identifier, parameters, device_interface=None):
    from bempp_amd.core.dense_assembler import DenseAssembler
    from bempp_amd.core.sparse_assembler import SparseAssembler

    if identifier in ("dense", "default_nonlocal"):
        return DenseAssembler(domain, dual_to_range, parameters)
    if identifier == "sparse":
        return SparseAssembler(domain, dual_to_range, parameters)
    raise ValueError("Unknown assembler type.")


class AssemblerInterface(object):
    """Binds spaces, an assembler, a device interface and a precision (assembler.py:42-97)."""

    def __init__(self, domain, dual_to_range, assembler, device_interface, precision, parameters=None):
        import bempp_amd.api as api

        self._domain = domain
        self._dual_to_range = dual_to_range
        self._parameters = api.assign_parameters(parameters)
        self._device_interface = device_interface if device_interface is not None else api.DEFAULT_DEVICE_INTERFACE
        self._precision = precision
        if not isinstance(assembler, str):
            self._implementation = assembler
        else:
            self._implementation = _create_assembler(domain, dual_to_range, assembler, self._parameters,
                                                     self._device_interface)

    @property
    def domain(self):
        return self._domain

    @property
    def dual_to_range(self):
        return self._dual_to_range

    @property
    def parameters(self):
        return self._parameters

    @property
    def device_interface(self):
        return self._device_interface

    def assemble(self, operator_descriptor, *args, **kwargs):
        return self._implementation.assemble(operator_descriptor, self._device_interface, self._precision, *args,
                                             **kwargs)


class PotentialAssembler(object):
    """Potential assembler interface (assembler.py:131-163)."""

    def __init__(self, space, points, operator_descriptor, device_interface, assembler, parameters):
        self.space = space
        self.points = points
        self.descriptor = operator_descriptor
        self.kernel_dimension = operator_descriptor.kernel_dimension
        self._is_complex = operator_descriptor.is_complex
        self._implementation = select_potential_implementation(
            space, points, operator_descriptor, device_interface, assembler, parameters
        )

    def evaluate(self, x):
        """Evaluate the potential for the coefficient vector x; shape (kernel_dimension, npoints)."""
        return self._implementation.evaluate(x)


def select_potential_implementation(space, points, operator_descriptor, device_interface, assembler, parameters):
    import bempp_amd.api as api

    parameters = api.assign_parameters(parameters)
    if device_interface is None:
        device_interface = api.DEFAULT_DEVICE_INTERFACE
    if assembler == "dense":
        from bempp_amd.core.dense_potential_assembler import DensePotentialAssembler

        return DensePotentialAssembler(space, operator_descriptor, points, device_interface, parameters)
    raise ValueError(f"Unknown potential assembler: {assembler}")
