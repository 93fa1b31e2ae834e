"""Generic boundary operator construction (bempp_cl/api/operators/boundary/common.py:10-64)."""


def create_operator(
    identifier,
    domain,
    range_,
    dual_to_range,
    parameters,
    assembler,
    operator_options,
    kernel_type,
    assembly_type,
    device_interface,
    precision,
    is_complex,
):
    """Create a boundary operator with an assembler and an OperatorDescriptor."""
    import bempp_amd.api as api
    from bempp_amd.api.assembly import assembler as _assembler
    from bempp_amd.api.assembly.boundary_operator import BoundaryOperatorWithAssembler
    from bempp_amd.api.operators import OperatorDescriptor

    if precision is None:
        precision = api.DEFAULT_PRECISION
    if precision not in ("single", "double"):
        raise ValueError("precision must be one of 'single' or 'double'")
    interface = _assembler.AssemblerInterface(domain, dual_to_range, assembler, device_interface, precision, parameters)
    descriptor = OperatorDescriptor(
        identifier, list(operator_options), kernel_type, assembly_type, precision, is_complex, None, 1
    )
    return BoundaryOperatorWithAssembler(domain, range_, dual_to_range, interface, descriptor)
