"""Direct solver on the device (bempp_cl/api/linalg/direct_solvers.py:6-60).

The reference factorises the host weak form with scipy.linalg (lu_factor / solve). Here the
dense weak form is already a device matrix and is factorised there (torch.linalg.lu_factor,
LU with partial pivoting in the matrix's dtype, as scipy does on the host); the right-hand side
is solved against the factors on the device.
"""

import numpy as _np


def compute_lu_factors(A):
    """LU factors (LU, pivots) of the weak form of A, as device tensors."""
    import torch

    return torch.linalg.lu_factor(A.weak_form().device_tensor())


def _solve(tensor, lu_factor, vec):
    import torch

    from bempp_amd.api.assembly.discrete_boundary_operator import _np_dtype

    dtype = _np.result_type(_np_dtype(tensor.dtype if lu_factor is None else lu_factor[0].dtype), vec.dtype)
    rhs = torch.as_tensor(_np.asarray(vec, dtype=dtype), device=tensor.device if lu_factor is None else lu_factor[0].device)
    if lu_factor is None:
        sol = torch.linalg.solve(tensor.to(rhs.dtype), rhs[:, None])
    else:
        lu, piv = lu_factor
        sol = torch.linalg.lu_solve(lu.to(rhs.dtype), piv, rhs[:, None])
    return sol[:, 0].cpu().numpy()


def lu(A, b, lu_factor=None):
    """Solve A x = b by LU decomposition; b a grid function (or a list for blocked operators)."""
    from bempp_amd.api.assembly.blocked_operator import (BlockedOperator, grid_function_list_from_coefficients,
                                                          projections_from_grid_functions_list)
    from bempp_amd.api.assembly.grid_function import GridFunction

    tensor = None if lu_factor is not None else A.weak_form().device_tensor()
    if isinstance(A, BlockedOperator):
        vec = projections_from_grid_functions_list(b, A.dual_to_range_spaces)
        return grid_function_list_from_coefficients(_solve(tensor, lu_factor, vec), A.domain_spaces)
    vec = b.projections(A.dual_to_range)
    return GridFunction(A.domain, coefficients=_solve(tensor, lu_factor, vec))
