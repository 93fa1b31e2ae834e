"""Sparse assembly of the L2 identity (bempp_cl/core/sparse_assembler.py:12-140).

The identity operator's weak form is the mass matrix M_ij = sum_e J_e sum_q w_q phi_i(q) psi_j(q)
over the elements in both supports, with the spaces' local multipliers, assembled once on the
host (bempp_cl/core/numba_kernels.py:713-800, 935-975). It is O(F) work, so it stays in
numpy/scipy: the dense operators it is combined with live on the GPU, and a sum of the two is
formed there (DenseDiscreteBoundaryOperator.__add__).
"""

import numpy as _np


def mass_matrix(domain, dual_to_range, order=4):
    """scipy CSR mass matrix (dual_to_range dofs x domain dofs) of two scalar spaces on one grid."""
    from scipy.sparse import coo_matrix

    from bempp_amd.api.integration.triangle_gauss import rule

    if domain.grid != dual_to_range.grid:
        raise ValueError("For sparse operators the domain and dual_to_range grids must be identical.")
    if domain.codomain_dimension != 1 or dual_to_range.codomain_dimension != 1:
        raise NotImplementedError("The L2 identity is provided for scalar spaces (DP0, DP1, P1).")
    points, weights = rule(order)
    elements = _np.flatnonzero(domain.support & dual_to_range.support)
    phi_t = dual_to_range.shapeset.evaluate(points)[0]  # (nt, nq)
    phi_s = domain.shapeset.evaluate(points)[0]  # (ns, nq)
    local = _np.einsum("iq,jq,q->ij", phi_t, phi_s, weights)
    nt, ns = local.shape
    jac = domain.grid.integration_elements[elements]
    mt = dual_to_range.local_multipliers[elements]
    ms = domain.local_multipliers[elements]
    vals = jac[:, None, None] * local[None, :, :] * mt[:, :, None] * ms[:, None, :]
    rows = _np.broadcast_to(dual_to_range.local2global[elements][:, :, None], vals.shape)
    cols = _np.broadcast_to(domain.local2global[elements][:, None, :], vals.shape)
    keep = vals != 0.0
    mat = coo_matrix(
        (vals[keep], (rows[keep].astype(_np.int64), cols[keep].astype(_np.int64))),
        shape=(dual_to_range.global_dof_count, domain.global_dof_count),
    )
    return mat.tocsr()


class SparseAssembler(object):
    """Assembler of sparse (local) operators: the L2 identity."""

    def __init__(self, domain, dual_to_range, parameters=None):
        import bempp_amd.api as api

        self.domain = domain
        self.dual_to_range = dual_to_range
        self.parameters = api.assign_parameters(parameters)

    def assemble(self, operator_descriptor, device_interface, precision, *args, **kwargs):
        from bempp_amd.api.assembly.discrete_boundary_operator import SparseDiscreteBoundaryOperator

        if operator_descriptor.assembly_type != "default_sparse" or operator_descriptor.kernel_type != "l2_identity":
            raise KeyError(f"Unsupported sparse operator {operator_descriptor.identifier}")
        return SparseDiscreteBoundaryOperator(
            mass_matrix(self.domain, self.dual_to_range, self.parameters.quadrature.regular))
