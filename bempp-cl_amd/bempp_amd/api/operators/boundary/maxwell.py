"""Maxwell electric field boundary operator (bempp_cl/api/operators/boundary/maxwell.py:8-38)."""

import numpy as _np

from bempp_amd.api.operators.boundary import common as _common


def electric_field(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                   device_interface=None, precision=None):
    """Electric field (EFIE) boundary operator: RWG domain, SNC dual_to_range."""
    if domain.identifier != "rwg0":
        raise ValueError("Domain space must be an RWG type function space.")
    if dual_to_range.identifier != "snc0":
        raise ValueError("Dual to range space must be an SNC type function space.")
    return _common.create_operator(
        "maxwell_electric_field_boundary", domain, range_, dual_to_range, parameters, assembler,
        [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], "helmholtz_single_layer",
        "maxwell_electric_field", device_interface, precision, True,
    )


def magnetic_field(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                   device_interface=None, precision=None):
    """Maxwell magnetic field (MFIE) boundary operator (bempp_cl/api/operators/boundary/maxwell.py:41-71).

    Assembled in FP64 in both precision modes (precision selects the output dtype).
    """
    if domain.identifier != "rwg0":
        raise ValueError("Domain space must be an RWG type function space.")
    if dual_to_range.identifier != "snc0":
        raise ValueError("Dual to range space must be an SNC type function space.")
    return _common.create_operator(
        "maxwell_magnetic_field_boundary", domain, range_, dual_to_range, parameters, assembler,
        [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], "helmholtz_single_layer", "maxwell_magnetic_field",
        device_interface, precision, True,
    )
