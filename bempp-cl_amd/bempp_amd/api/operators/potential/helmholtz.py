"""Helmholtz potential operators (bempp_cl/api/operators/potential/helmholtz.py:6-100)."""

import numpy as _np


def _make(identifier, kernel, space, points, wavenumber, parameters, assembler, device_interface, precision):
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if _np.real(wavenumber) == 0:
        raise ValueError("Purely imaginary wavenumbers (modified Helmholtz potentials) are not supported.")
    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(
        identifier, [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], kernel, "default_scalar", precision,
        True, None, 1,
    )
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))


def single_layer(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Helmholtz single-layer potential."""
    return _make("helmholtz_single_layer_potential", "helmholtz_single_layer", space, points, wavenumber, parameters,
                 assembler, device_interface, precision)


def double_layer(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Helmholtz double-layer potential."""
    return _make("helmholtz_double_layer_potential", "helmholtz_double_layer", space, points, wavenumber, parameters,
                 assembler, device_interface, precision)
