"""Laplace potential operators (bempp_cl/api/operators/potential/laplace.py)."""


def _make(identifier, kernel, space, points, parameters, assembler, device_interface, precision):
    import bempp_amd.api as api
    from bempp_amd.api.assembly.assembler import PotentialAssembler
    from bempp_amd.api.assembly.potential_operator import PotentialOperator
    from bempp_amd.api.operators import OperatorDescriptor

    if precision is None:
        precision = api.DEFAULT_PRECISION
    descriptor = OperatorDescriptor(identifier, [], kernel, "default_scalar", precision, False, None, 1)
    return PotentialOperator(PotentialAssembler(space, points, descriptor, device_interface, assembler, parameters))


def single_layer(space, points, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Laplace single-layer potential."""
    return _make("laplace_single_layer_potential", "laplace_single_layer", space, points, parameters, assembler,
                 device_interface, precision)


def double_layer(space, points, parameters=None, assembler="dense", device_interface=None, precision=None):
    """Laplace double-layer potential."""
    return _make("laplace_double_layer_potential", "laplace_double_layer", space, points, parameters, assembler,
                 device_interface, precision)
