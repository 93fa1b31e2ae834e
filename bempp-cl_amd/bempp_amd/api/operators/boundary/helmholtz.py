"""Helmholtz boundary operators (bempp_cl/api/operators/boundary/helmholtz.py:9-178).

The options array is [Re k, Im k] (helmholtz.py:41). The reference redirects Re k == 0 to
modified Helmholtz operators (helmholtz.py:22-32), which this library does not provide.
"""

import numpy as _np

from bempp_amd.api.operators.boundary import common as _common


def _check_wavenumber(wavenumber):
    if _np.real(wavenumber) == 0:
        raise ValueError("Purely imaginary wavenumbers (modified Helmholtz operators) are not supported.")


def _make(identifier, kernel, domain, range_, dual_to_range, wavenumber, parameters, assembler, device_interface,
          precision, assembly_type="default_scalar"):
    _check_wavenumber(wavenumber)
    return _common.create_operator(
        identifier, domain, range_, dual_to_range, parameters, assembler,
        [float(_np.real(wavenumber)), float(_np.imag(wavenumber))], kernel, assembly_type,
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


def hypersingular(domain, range_, dual_to_range, wavenumber, parameters=None, assembler="default_nonlocal",
                  device_interface=None, precision=None):
    """Helmholtz hypersingular boundary operator (bempp_cl/api/operators/boundary/helmholtz.py:134-178)."""
    if domain.shapeset.identifier != "p1_discontinuous":
        raise ValueError("Domain shapeset must be of type 'p1_discontinuous'.")
    if dual_to_range.shapeset.identifier != "p1_discontinuous":
        raise ValueError("Dual to range shapeset must be of type 'p1_discontinuous'.")
    return _make("helmholtz_hypersingular_boundary", "helmholtz_single_layer", domain, range_, dual_to_range,
                 wavenumber, parameters, assembler, device_interface, precision, "helmholtz_hypersingular")


def single_layer_frequency_sweep(domain, range_, dual_to_range, wavenumbers, parameters=None, device_interface=None,
                                 precision=None):
    """Weak forms of the Helmholtz single layer at many real wavenumbers (DP0 x DP0 spaces).

    Returns one DenseDiscreteBoundaryOperator per wavenumber, each equal to
    `single_layer(domain, range_, dual_to_range, k, assembler="dense").weak_form()`, assembled
    together by the frequency-batched device kernel. The matrices are views of one
    (len(wavenumbers), rows, cols) device tensor.
    """
    import bempp_amd.api as api
    from bempp_amd.api.assembly.discrete_boundary_operator import DenseDiscreteBoundaryOperator
    from bempp_amd.core.dispatcher import check_device_interface
    from bempp_amd.core.frequency_sweep import hip_kbatch_assemble

    for k in _np.atleast_1d(wavenumbers):
        _check_wavenumber(k)
    check_device_interface(device_interface or api.DEFAULT_DEVICE_INTERFACE)
    mats = hip_kbatch_assemble(domain, dual_to_range, wavenumbers, precision or api.DEFAULT_PRECISION,
                               parameters=parameters)
    return [DenseDiscreteBoundaryOperator(mats[m]) for m in range(mats.shape[0])]
