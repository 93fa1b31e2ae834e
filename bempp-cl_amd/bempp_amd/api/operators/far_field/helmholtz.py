"""Helmholtz far-field operators (bempp_cl/api/operators/far_field/helmholtz.py:6-76).

The points are unit directions; the kernel is e^{-i Re(k) x.y} / (4 pi) as in the reference
(bempp_cl/core/numba_kernels.py:402-449). Evaluated in FP64 on the device; `precision` selects
the output dtype (complex64 for "single").
"""

import numpy as _np


def _far_field(identifier, kernel, space, points, wavenumber, parameters, assembler, device_interface, precision):
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(
        identifier, [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], kernel, "default_scalar", precision,
        True, None, 1,
    )
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))


def single_layer(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Helmholtz single-layer far-field operator."""
    return _far_field("helmholtz_far_field_single_layer_potential", "helmholtz_far_field_single_layer", space, points,
                      wavenumber, parameters, assembler, device_interface, precision)


def double_layer(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Helmholtz double-layer far-field operator."""
    return _far_field("helmholtz_far_field_double_layer_potential", "helmholtz_far_field_double_layer", space, points,
                      wavenumber, parameters, assembler, device_interface, precision)
