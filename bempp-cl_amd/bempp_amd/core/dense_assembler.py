"""Dense assembly of boundary operators on the device (bempp_cl/core/dense_assembler.py:7-100).

`DenseAssembler.assemble` allocates the (rows, cols) result in HBM, runs the regular-pair
kernel over all test x trial support elements and, when the grids coincide, the FP64
singular kernel over the Sauter-Schwab pairs, which adds its values straight into the
result with device atomics (the reference's np.add.at step, dense_assembler.py:78-98).

Precision: "double" runs everything in FP64 and returns float64/complex128; "single"
runs the MIXED mode (FP32 far pairs, FP64 near and singular pairs) and returns
float32/complex64.
"""

from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR
import numpy as _np

from bempp_amd.api.utils.helpers import get_type


class DenseAssembler(object):
    """Dense assembler of integral operators (dense_assembler.py:7-46)."""

    def __init__(self, domain, dual_to_range, parameters=None):
        import bempp_amd.api as api

        self._domain = domain
        self._dual_to_range = dual_to_range
        self._parameters = api.assign_parameters(parameters)

    @property
    def domain(self):
        return self._domain

    @property
    def dual_to_range(self):
        return self._dual_to_range

    @property
    def parameters(self):
        return self._parameters

    def assemble(self, operator_descriptor, device_interface, precision, *args, **kwargs):
        from bempp_amd.api.assembly.discrete_boundary_operator import DenseDiscreteBoundaryOperator
        from bempp_amd.core.dispatcher import dense_assembler_dispatcher

        if self.domain.requires_dof_transformation or self.dual_to_range.requires_dof_transformation:
            raise ValueError("Spaces that require dof transformations not supported for dense assembly.")
        mat = dense_assembler_dispatcher(
            device_interface, self.domain, self.dual_to_range, self.parameters, operator_descriptor
        )
        if self.parameters.assembly.always_promote_to_double:
            import torch

            if mat.dtype == torch.float32:
                mat = mat.double()
            elif mat.dtype == torch.complex64:
                mat = mat.to(torch.complex128)
        return DenseDiscreteBoundaryOperator(mat)


def _kernel_name(descriptor):
    if descriptor.assembly_type in ("maxwell_electric_field", "laplace_hypersingular", "helmholtz_hypersingular",
                                    "maxwell_electric_far_field", "maxwell_magnetic_far_field", "maxwell_magnetic_field"):
        return descriptor.assembly_type
    if descriptor.assembly_type != "default_scalar":
        raise KeyError(f"Unsupported assembly type {descriptor.assembly_type}")
    return descriptor.kernel_type


def _wavenumber(descriptor):
    opts = list(descriptor.options)
    if len(opts) == 0:
        return 0.0
    return complex(opts[0], opts[1] if len(opts) > 1 else 0.0)


def _is_localised_injective(space):
    """True if every (support element, shape function) owns its own dof (DP0/DP1)."""
    return space.is_localised


def hip_dense_assemble(domain, dual_to_range, parameters, operator_descriptor, out=None, singular=True,
                       accumulate=True, events=None):
    """Assemble the dense matrix into a new (or the given) device tensor and return it.

    out: optional (rows, cols) device tensor; the operator is added to it when accumulate is
    true, otherwise it is overwritten. events: optional dict of torch.cuda.Event objects with
    keys "regular_start", "regular_end", "singular_end", recorded around the launches on the
    current stream (used for kernel timing).
    """
    import torch
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.singular_assembler import singular_rule

    op = _kernel_name(operator_descriptor)
    if op not in hb.OPERATORS:
        raise KeyError(f"Unsupported kernel {op}")
    precision = operator_descriptor.precision
    types = get_type(precision)
    dtype = types.complex if operator_descriptor.is_complex else types.real
    tdtype = getattr(torch, dtype)
    rows, cols = dual_to_range.global_dof_count, domain.global_dof_count
    dev = hb.device()
    grids_identical = domain.grid == dual_to_range.grid

    test = hb.device_space(dual_to_range)
    trial = hb.device_space(domain)
    points, weights = triangle_rule(parameters.quadrature.regular)
    rule = hb.quad_rule(points, weights)
    mode = hb.PRECISION_MIXED if precision == "single" else hb.PRECISION_FP64
    # rules of more than 16 points (regular orders 7-20) take the library's one-lane-per-pair FP64
    # kernel in both precisions (stored in the precision's dtype): no near-field split, no
    # symmetric traversal
    wide = len(weights) > hb.QUAD_FAST_POINTS
    if wide:
        if op == "maxwell_magnetic_field":
            raise ValueError(f"the magnetic field operator takes regular rules of at most {hb.QUAD_FAST_POINTS} "
                             f"points (order <= 8); got order {parameters.quadrature.regular}")
        mode = hb.PRECISION_FP64
    direct = _is_localised_injective(domain) and _is_localised_injective(dual_to_range)
    full_cover = (
        direct
        and dual_to_range.number_of_support_elements * dual_to_range.number_of_shape_functions == rows
        and domain.number_of_support_elements * domain.number_of_shape_functions == cols
    )

    if out is None:
        out = torch.empty((rows, cols), dtype=tdtype, device=dev)
        accumulate = False
    else:
        if tuple(out.shape) != (rows, cols) or out.dtype != tdtype or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous ({rows}, {cols}) {dtype} tensor")
    hip_opts = getattr(parameters, "hip", None)
    sym_opt = getattr(hip_opts, "symmetric", True) if hip_opts is not None else True
    # a single layer or hypersingular operator on one P1 space: every unordered pair once; the
    # library zeroes the output and adds the transpose itself
    sym_sl = (not accumulate and not wide and sym_opt and op in _SYMMETRIC_P1_OPERATORS
              and use_symmetric_pass(domain, dual_to_range))
    # the electric field operator on one RWG dof map (RWG trial, SNC or RWG test) is symmetric too
    sym_efie = (not accumulate and not wide and sym_opt and op == "maxwell_electric_field"
                and use_symmetric_efie(domain, dual_to_range))
    sym_sl = sym_sl or sym_efie
    if not accumulate and not full_cover and not sym_sl:
        hb.fill_zero(out)
    near = getattr(hip_opts, "mixed_near_field_factor", DEFAULT_NEAR_FIELD_FACTOR) if hip_opts is not None else DEFAULT_NEAR_FIELD_FACTOR
    kahan = getattr(hip_opts, "kahan", False) if hip_opts is not None else False
    srule = None
    if grids_identical and singular:
        srule = singular_rule(dual_to_range.grid, parameters.quadrature.singular, dual_to_range.support, domain.support)

    hip_opts_ = getattr(parameters, "hip", None)
    upper = (
        direct
        and not wide
        and not accumulate
        and use_upper_triangle(domain, dual_to_range, op)
        and (getattr(hip_opts_, "symmetric", True) if hip_opts_ is not None else True)
    )
    if events is not None:
        events["regular_start"].record()
    # the magnetic field operator adds to the (zeroed) output: FP64 without dof blocks; the mixed mode
    # takes the RWG dof blocks (far pairs in FP32) and the near pass below
    mfie = op == "maxwell_magnetic_field"
    if mfie and mode != hb.PRECISION_MIXED:
        hb.regular_assemble(op, test, trial, grids_identical, rule, _wavenumber(operator_descriptor), mode, out,
                            accumulate=1, near_factor=near, kahan=kahan)
    elif upper:
        hb.regular_assemble_symmetric(op, test, rule, _wavenumber(operator_descriptor), mode, out, near_factor=near,
                                      kahan=kahan)
    elif sym_sl:
        from bempp_amd.core.blocking import device_symmetric_blocks

        role = "trial_rwg_mixed" if sym_efie and mode == hb.PRECISION_MIXED else "trial"
        hb.regular_assemble_sl_dl_symmetric(op, device_symmetric_blocks(domain, role), rule,
                                            _wavenumber(operator_descriptor), mode, out, None, near_factor=near,
                                            kahan=kahan)
    elif direct:
        hb.regular_assemble(
            op, test, trial, grids_identical, rule, _wavenumber(operator_descriptor), mode, out,
            accumulate=int(accumulate), near_factor=near, kahan=kahan,
        )
    else:
        from bempp_amd.core.blocking import device_blocks

        # the mixed EFIE / MFIE kernels take 256-element trial blocks of up to ~400 edges
        rwg_mixed = op in ("maxwell_electric_field", "maxwell_magnetic_field") and mode == hb.PRECISION_MIXED
        hb.regular_assemble_blocked(
            op, device_blocks(dual_to_range, "test_rwg_mixed" if rwg_mixed else "test"),
            device_blocks(domain, "trial_rwg_mixed" if rwg_mixed else "trial"),
            grids_identical, rule, _wavenumber(operator_descriptor), mode, out, near_factor=near, kahan=kahan,
        )
    if events is not None:
        events["regular_end"].record()
    if mode == hb.PRECISION_MIXED:
        from bempp_amd.core.near_field import device_near_candidates

        cand = device_near_candidates(dual_to_range, domain, near)
        hb.near_assemble(op, test, trial, cand, grids_identical, rule, _wavenumber(operator_descriptor), near, out)
    if events is not None and "near_end" in events:
        events["near_end"].record()
    if srule is not None:
        hb.singular_assemble(
            op, test, trial, srule.pairs, srule.test_points, srule.trial_points, srule.weights,
            _wavenumber(operator_descriptor), out,
        )
    if events is not None:
        events["singular_end"].record()
    return out


_SL_DL = {
    ("laplace_single_layer", "laplace_double_layer"),
    ("helmholtz_single_layer", "helmholtz_double_layer"),
}


def can_fuse_sl_dl(first, second):
    """True if the two (descriptor, domain, dual_to_range) triples form a fusable single/double pair."""
    d1, dom1, dual1 = first
    d2, dom2, dual2 = second
    return (
        (d1.assembly_type, d2.assembly_type) == ("default_scalar", "default_scalar")
        and (d1.kernel_type, d2.kernel_type) in _SL_DL
        and tuple(d1.options) == tuple(d2.options)
        and d1.precision == d2.precision
        and dom1 == dom2
        and dual1 == dual2
    )


_SYMMETRIC_SINGLE_LAYERS = ("laplace_single_layer", "helmholtz_single_layer")
# operators whose matrix on one space is symmetric (the symmetric P1 traversal)
_SYMMETRIC_P1_OPERATORS = _SYMMETRIC_SINGLE_LAYERS + ("laplace_hypersingular", "helmholtz_hypersingular")


def use_upper_triangle(domain, dual_to_range, op):
    """True if a single layer on one DP0 space can be assembled as its upper triangle plus a
    mirror copy: the operator is symmetric and the dof of list position p is p."""
    if op not in _SYMMETRIC_SINGLE_LAYERS or domain != dual_to_range or domain.number_of_shape_functions != 1:
        return False
    if not _is_localised_injective(domain):
        return False
    n = domain.number_of_support_elements
    dofs = domain.local2global[domain.support_elements, 0]
    return domain.global_dof_count == n and bool(_np.array_equal(dofs, _np.arange(n)))


def use_symmetric_pass(domain, dual_to_range):
    """True if the fused pair can run the symmetric pass: one P1 space on both sides."""
    return (
        domain == dual_to_range
        and not _is_localised_injective(domain)
        and domain.codomain_dimension == 1
        and domain.number_of_shape_functions == 3
    )


def use_symmetric_efie(domain, dual_to_range):
    """True if the electric field operator can run the symmetric pass: both spaces carry RWG
    or SNC functions (the kernel evaluates the SNC test side through the RWG functions) over the
    same support, dof map, multipliers and normal orientation, so entry (i, j) equals (j, i) up
    to the singular quadrature."""
    kinds = ("rwg0", "snc0")
    if domain.shapeset.identifier not in kinds or dual_to_range.shapeset.identifier not in kinds:
        return False
    if domain.grid != dual_to_range.grid or domain.global_dof_count != dual_to_range.global_dof_count:
        return False
    return bool(
        _np.array_equal(domain.support_elements, dual_to_range.support_elements)
        and _np.array_equal(domain.local2global, dual_to_range.local2global)
        and _np.array_equal(domain.local_multipliers, dual_to_range.local_multipliers)
        and _np.array_equal(domain.normal_multipliers, dual_to_range.normal_multipliers)
    )


def hip_dense_assemble_sl_dl(domain, dual_to_range, parameters, sl_descriptor, dl_descriptor, out_single=None,
                             out_double=None, events=None):
    """Assemble a single layer and the double layer of the same family in one pass.

    Both matrices equal those of two hip_dense_assemble calls (the same quadrature and
    precision); every point pair's distance, 1/r and e^{ikr} are computed once for both.
    Returns (single layer, double layer) device tensors. events: optional dict of
    torch.cuda.Event with keys "regular_start", "regular_end", "singular_end".
    """
    import torch
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.singular_assembler import singular_rule

    if not can_fuse_sl_dl((sl_descriptor, domain, dual_to_range), (dl_descriptor, domain, dual_to_range)):
        raise ValueError("expected a single layer and the double layer of the same family, wavenumber and precision")
    op = _kernel_name(sl_descriptor)
    precision = sl_descriptor.precision
    types = get_type(precision)
    dtype = types.complex if sl_descriptor.is_complex else types.real
    tdtype = getattr(torch, dtype)
    rows, cols = dual_to_range.global_dof_count, domain.global_dof_count
    dev = hb.device()
    grids_identical = domain.grid == dual_to_range.grid
    test = hb.device_space(dual_to_range)
    trial = hb.device_space(domain)
    points, weights = triangle_rule(parameters.quadrature.regular)
    if len(weights) > hb.QUAD_FAST_POINTS:
        # rules of more than 16 points have no fused pass: one assembly per operator
        results = []
        for desc, o in ((sl_descriptor, out_single), (dl_descriptor, out_double)):
            results.append(hip_dense_assemble(domain, dual_to_range, parameters, desc, out=o, accumulate=False))
        return results[0], results[1]
    rule = hb.quad_rule(points, weights)
    mode = hb.PRECISION_MIXED if precision == "single" else hb.PRECISION_FP64
    direct = _is_localised_injective(domain) and _is_localised_injective(dual_to_range)
    full_cover = (
        direct
        and dual_to_range.number_of_support_elements * dual_to_range.number_of_shape_functions == rows
        and domain.number_of_support_elements * domain.number_of_shape_functions == cols
    )
    hip_opts = getattr(parameters, "hip", None)
    near = getattr(hip_opts, "mixed_near_field_factor", DEFAULT_NEAR_FIELD_FACTOR) if hip_opts is not None else DEFAULT_NEAR_FIELD_FACTOR
    kahan = getattr(hip_opts, "kahan", False) if hip_opts is not None else False
    symmetric = (
        getattr(hip_opts, "symmetric", True) if hip_opts is not None else True
    ) and use_symmetric_pass(domain, dual_to_range)
    k = _wavenumber(sl_descriptor)
    outs = []
    for o in (out_single, out_double):
        if o is None:
            o = torch.empty((rows, cols), dtype=tdtype, device=dev)
        elif tuple(o.shape) != (rows, cols) or o.dtype != tdtype or not o.is_contiguous():
            raise ValueError(f"outputs must be contiguous ({rows}, {cols}) {dtype} tensors")
        if not full_cover and not symmetric:
            hb.fill_zero(o)  # the symmetric pass overwrites its outputs itself
        outs.append(o)

    if events is not None:
        events["regular_start"].record()
    if symmetric:
        from bempp_amd.core.blocking import device_symmetric_blocks

        hb.regular_assemble_sl_dl_symmetric(op, device_symmetric_blocks(domain), rule, k, mode, outs[0], outs[1],
                                            near_factor=near, kahan=kahan)
    elif direct:
        hb.regular_assemble_sl_dl(op, test, trial, None, None, grids_identical, rule, k, mode, outs[0], outs[1],
                                  accumulate=0, near_factor=near, kahan=kahan)
    else:
        from bempp_amd.core.blocking import device_blocks

        hb.regular_assemble_sl_dl(op, test, trial, device_blocks(dual_to_range, "test_pair"),
                                  device_blocks(domain, "trial"),
                                  grids_identical, rule, k, mode, outs[0], outs[1], near_factor=near, kahan=kahan)
    if events is not None:
        events["regular_end"].record()
    if mode == hb.PRECISION_MIXED:
        from bempp_amd.core.near_field import device_near_candidates

        if symmetric:
            cand = device_near_candidates(dual_to_range, domain, near, unordered=True)
            hb.near_assemble_sl_dl_symmetric(op, test, cand, rule, k, near, outs[0], outs[1])
        else:
            cand = device_near_candidates(dual_to_range, domain, near)
            hb.near_assemble_sl_dl(op, test, trial, cand, grids_identical, rule, k, near, outs[0], outs[1])
    if grids_identical:
        srule = singular_rule(dual_to_range.grid, parameters.quadrature.singular, dual_to_range.support, domain.support)
        hb.singular_assemble_sl_dl(op, test, trial, srule.pairs, srule.test_points, srule.trial_points, srule.weights,
                                   k, outs[0], outs[1])
    if events is not None:
        events["singular_end"].record()
    return outs[0], outs[1]


def assemble_dense(domain, dual_to_range, parameters, operator_descriptor, device_interface):
    """Assemble and return the dense matrix as a numpy array (dense_assembler.py:49-100)."""
    from bempp_amd.core.dispatcher import dense_assembler_dispatcher

    mat = dense_assembler_dispatcher(device_interface, domain, dual_to_range, parameters, operator_descriptor)
    return mat.cpu().numpy()
