"""Helmholtz boundary operators (bempp_cl/api/operators/boundary/helmholtz.py:9-131).

The options array is [Re k, Im k] (helmholtz.py:41). The reference redirects Re k == 0 to
modified Helmholtz operators (helmholtz.py:22-32), which this library does not provide.
"""

import numpy as _np

from bempp_amd.api.operators.boundary import common as _common


def _check_wavenumber(wavenumber):
    if _np.real(wavenumber) == 0:
        raise ValueError("Purely imaginary wavenumbers (modified Helmholtz operators) are not supported.")


def _make(identifier, kernel, domain, range_, dual_to_range, wavenumber, parameters, assembler, device_interface,
          precision):
    _check_wavenumber(wavenumber)
    return _common.create_operator(
        identifier, domain, range_, dual_to_range, parameters, assembler,
        [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], kernel, "default_scalar",
        device_interface, precision, True,
    )


def single_layer(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                 device_interface=None, precision=None):
    """Helmholtz single-layer boundary operator."""
    return _make("helmholtz_single_layer_boundary", "helmholtz_single_layer", domain, range_, dual_to_range,
                 wavenumber, parameters, assembler, device_interface, precision)


def double_layer(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                 device_interface=None, precision=None):
    """Helmholtz double-layer boundary operator."""
    return _make("helmholtz_double_layer_boundary", "helmholtz_double_layer", domain, range_, dual_to_range,
                 wavenumber, parameters, assembler, device_interface, precision)


def adjoint_double_layer(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                         device_interface=None, precision=None):
    """Helmholtz adjoint double-layer boundary operator."""
    return _make("helmholtz_adjoint_double_layer_boundary", "helmholtz_adjoint_double_layer", domain, range_,
                 dual_to_range, wavenumber, parameters, assembler, device_interface, precision)
