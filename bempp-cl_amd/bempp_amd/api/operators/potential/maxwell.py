"""Maxwell electric and magnetic field potentials (bempp_cl/api/operators/potential/maxwell.py:6-75)."""

import numpy as _np


def electric_field(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None,
                   precision=None):
    """Electric field potential of an RWG density; kernel dimension 3."""
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if space.identifier != "rwg0":
        raise ValueError("Space must be an RWG type function space.")
    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(
        "maxwell_electric_field_potential", [float(_np.real(wavenumber)), float(_np.imag(wavenumber))],
        "helmholtz_single_layer", "maxwell_electric_field", precision, True, None, 3,
    )
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))


def magnetic_field(space, points, wavenumber, parameters=None, assembler="dense", device_interface=None,
                   precision=None):
    """Magnetic field potential of an RWG density; kernel dimension 3 (evaluated in FP64)."""
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if space.identifier != "rwg0":
        raise ValueError("Space must be an RWG type function space.")
    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(
        "maxwell_magnetic_field_potential", [float(_np.real(wavenumber)), float(_np.imag(wavenumber))],
        "helmholtz_single_layer", "maxwell_magnetic_field", precision, True, None, 3,
    )
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))
