"""Dense potential and far-field evaluation on the device (bempp_cl/core/dense_potential_assembler.py:4-40).

The evaluator maps space coefficients to the full grid (space.map_to_full_grid @ x, which
applies the local multipliers, space.py:376-391), uploads them and the evaluation points,
and runs bem_potential_evaluate. The result has shape (kernel_dimension, npoints).
"""

import numpy as _np


class DensePotentialAssembler(object):
    """Implementation of a dense potential assembler."""

    def __init__(self, space, operator_descriptor, points, device_interface, parameters=None):
        from bempp_amd.core.dispatcher import potential_dispatcher

        self.space = space
        self._implementation = potential_dispatcher(
            device_interface, space, operator_descriptor, points, parameters
        )

    def evaluate(self, x):
        return self._implementation(x)


def hip_potential_assembler(space, operator_descriptor, points, parameters):
    import torch
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import _kernel_name, _wavenumber

    op = _kernel_name(operator_descriptor)
    if op not in ("laplace_single_layer", "laplace_double_layer", "helmholtz_single_layer",
                  "helmholtz_double_layer", "maxwell_electric_field", "helmholtz_far_field_single_layer",
                  "helmholtz_far_field_double_layer", "maxwell_electric_far_field", "maxwell_magnetic_far_field",
                  "maxwell_magnetic_field"):
        raise KeyError(f"Unsupported potential kernel {op}")
    dev = hb.device()
    dspace = hb.device_space(space)
    pts = _np.asarray(points, dtype=_np.float64)
    if pts.ndim != 2 or pts.shape[0] != 3:
        raise ValueError("points must have shape (3, npoints)")
    dpts = torch.from_numpy(_np.ascontiguousarray(pts.T)).to(dev)
    qp, qw = triangle_rule(parameters.quadrature.regular)
    rule = hb.quad_rule(qp, qw)
    mode = hb.PRECISION_MIXED if operator_descriptor.precision == "single" else hb.PRECISION_FP64
    kdim = operator_descriptor.kernel_dimension
    out_dtype = torch.complex64 if operator_descriptor.precision == "single" else torch.complex128
    k = _wavenumber(operator_descriptor)
    full_map = (space.map_to_full_grid @ space.dof_transformation).tocsr()
    # every (element, shape function) row of the map holds at most one entry (a dof and its
    # multiplier): the coefficients are then mapped on the device by a gather
    gather = None
    if full_map.shape[0] == 0 or _np.diff(full_map.indptr).max() <= 1:
        nnz = _np.diff(full_map.indptr)
        col = _np.zeros(full_map.shape[0], dtype=_np.int64)
        val = _np.zeros(full_map.shape[0], dtype=_np.complex128)
        rows = _np.nonzero(nnz)[0]
        col[rows] = full_map.indices[full_map.indptr[rows]]
        val[rows] = full_map.data[full_map.indptr[rows]]
        gather = (torch.from_numpy(col).to(dev), torch.from_numpy(val).to(dev))

    def evaluator(x, as_tensor=False):
        if torch.is_tensor(x):
            coeffs = x
        elif gather is not None:
            xh = _np.ascontiguousarray(_np.asarray(x), dtype=_np.complex128)
            if xh.ndim != 1 or xh.shape[0] != full_map.shape[1]:
                raise ValueError(f"expected {full_map.shape[1]} coefficients, got an array of shape {xh.shape}")
            xd = torch.from_numpy(xh).to(dev)
            coeffs = gather[1] * xd[gather[0]] if xd.numel() else torch.zeros_like(gather[1])
        else:
            xf = full_map @ _np.asarray(x)
            coeffs = torch.from_numpy(_np.ascontiguousarray(xf.astype(_np.complex128))).to(dev)
        out = torch.empty((kdim, pts.shape[1]), dtype=out_dtype, device=dev)
        ws = hb.potential_evaluate(op, dspace, coeffs, dpts, rule, k, mode, out)
        if as_tensor:
            return out, ws
        # through page-locked memory from torch's caching host allocator: a pageable copy is
        # staged by the runtime and takes several times longer for the 1.5 MB of a 65,536-point
        # vector potential
        host = torch.empty(out.shape, dtype=out.dtype, pin_memory=True)
        host.copy_(out)
        res = host.numpy()
        if not operator_descriptor.is_complex:
            res = res.real.copy() if not _np.iscomplexobj(x) else res
        return res

    return evaluator
