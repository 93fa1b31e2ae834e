"""Sparse boundary operators (bempp_cl/api/operators/boundary/sparse.py:6-26)."""

from bempp_amd.api.operators.boundary import common as _common


def identity(domain, range_, dual_to_range, parameters=None, device_interface=None, precision=None):
    """Assemble the L^2 identity operator (the mass matrix of the two spaces)."""
    return _common.create_operator(
        "l2_identity", domain, range_, dual_to_range, parameters, "sparse", [], "l2_identity", "default_sparse",
        device_interface, precision, False,
    )
