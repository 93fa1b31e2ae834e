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
    full_map = space.map_to_full_grid @ space.dof_transformation

    def evaluator(x, as_tensor=False):
        if torch.is_tensor(x):
            coeffs = x
        else:
            xf = full_map @ _np.asarray(x)
            coeffs = torch.from_numpy(_np.ascontiguousarray(xf.astype(_np.complex128))).to(dev)
        out = torch.empty((kdim, pts.shape[1]), dtype=out_dtype, device=dev)
        ws = hb.potential_evaluate(op, dspace, coeffs, dpts, rule, k, mode, out)
        if as_tensor:
            return out, ws
        res = out.cpu().numpy()
        if not operator_descriptor.is_complex:
            res = res.real.copy() if not _np.iscomplexobj(x) else res
        return res

    return evaluator
