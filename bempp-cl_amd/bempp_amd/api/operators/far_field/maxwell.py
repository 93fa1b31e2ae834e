"""Maxwell far-field operators (bempp_cl/api/operators/far_field/maxwell.py:6-70).

Electric and magnetic far fields of an RWG density at unit directions (kernel dimension 3),
bempp_cl/core/numba_kernels.py:2437-2570. Evaluated in FP64 on the device; `precision` selects
the output dtype.
"""

import numpy as _np


def _far_field(identifier, assembly_type, space, points, wavenumber, parameters, assembler, device_interface,
               precision):
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if space.identifier != "rwg0":
        raise ValueError("Space must be an RWG type function space.")
    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(
        identifier, [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], "helmholtz_far_field_single_layer",
        assembly_type, precision, True, None, 3,
    )
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))


def electric_field(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None,
                   precision=None):
    """Maxwell electric far-field operator."""
    return _far_field("maxwell_far_field_electric_field_potential", "maxwell_electric_far_field", space, points,
                      wavenumber, parameters, assembler, device_interface, precision)


def magnetic_field(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None,
                   precision=None):
    """Maxwell magnetic far-field operator."""
    return _far_field("maxwell_far_field_magnetic_field_potential", "maxwell_magnetic_far_field", space, points,
                      wavenumber, parameters, assembler, device_interface, precision)
