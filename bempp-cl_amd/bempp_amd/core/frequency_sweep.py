"""Frequency sweep of the Helmholtz single layer on the device.

The reference assembles a sweep as one `helmholtz.single_layer(..., k).weak_form()` call per
wavenumber (bempp_cl/api/operators/boundary/helmholtz.py:14-48, each going through
bempp_cl/core/dense_assembler.py:49-100). Here all wavenumbers of a DP0 x DP0 sweep are
assembled together: one persistent launch for the regular pairs (the point geometry of a
pair is computed once and reused for every wavenumber), then the FP64 near-field pairs of
the mixed mode and the FP64 Sauter-Schwab singular pairs for all wavenumbers. The launch
shape comes from the autotuner (bempp_amd/core/autotune.py).
"""

from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR
import numpy as _np


def hip_kbatch_assemble(domain, dual_to_range, wavenumbers, precision="single", parameters=None, out=None,
                        config=None, autotune=True, events=None):
    """Assemble the Helmholtz single layer for every real wavenumber into a (nk, rows, cols) tensor.

    precision: "single" (mixed FP32/FP64, complex64) or "double" (FP64, complex128).
    config: explicit (test_tile, workgroups_per_cu); otherwise the tuned (or default) one.
    events: optional dict of torch.cuda.Event with keys "regular_start", "regular_end", "end"; the
    regular launch and the near/singular launches are then issued as two calls and timed apart.
    """
    import torch
    import bempp_amd.api as api
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import autotune as at
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.near_field import device_near_candidates
    from bempp_amd.core.singular_assembler import singular_rule

    parameters = api.assign_parameters(parameters)
    ks = _np.asarray(wavenumbers).ravel()
    if _np.iscomplexobj(ks) or not _np.issubdtype(ks.dtype, _np.number) and ks.size:
        raise ValueError("the frequency sweep takes finite real wavenumbers")
    ks = ks.astype(_np.float64)
    if ks.size == 0 or ks.size > hb.KBATCH_MAX_WAVENUMBERS:
        raise ValueError(f"between 1 and {hb.KBATCH_MAX_WAVENUMBERS} wavenumbers are supported")
    if not _np.all(_np.isfinite(ks)):
        raise ValueError("the frequency sweep takes finite real wavenumbers")
    if _np.any(ks == 0):
        raise ValueError("Wavenumber must not be zero (use the Laplace single layer).")
    for sp in (domain, dual_to_range):
        if sp.shapeset.identifier != "p0_discontinuous":
            raise ValueError("the frequency sweep supports DP0 spaces")
    if precision not in ("single", "double"):
        raise ValueError("precision must be 'single' or 'double'")
    mode = hb.PRECISION_MIXED if precision == "single" else hb.PRECISION_FP64
    tdtype = torch.complex64 if precision == "single" else torch.complex128
    rows, cols = dual_to_range.global_dof_count, domain.global_dof_count
    nk = ks.size
    dev = hb.device()
    if out is None:
        out = torch.empty((nk, rows, cols), dtype=tdtype, device=dev)
    elif tuple(out.shape) != (nk, rows, cols) or out.dtype != tdtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous ({nk}, {rows}, {cols}) {tdtype} tensor")
    covers = (dual_to_range.number_of_support_elements == rows and domain.number_of_support_elements == cols)
    if not covers:
        hb.fill_zero(out)

    grids_identical = domain.grid == dual_to_range.grid
    test = hb.device_space(dual_to_range)
    trial = hb.device_space(domain)
    points, weights = triangle_rule(parameters.quadrature.regular)
    if len(weights) != 6:
        raise ValueError("the frequency sweep uses the order-4 (6-point) regular rule")
    rule = hb.quad_rule(points, weights)
    hip_opts = getattr(parameters, "hip", None)
    near = getattr(hip_opts, "mixed_near_field_factor", DEFAULT_NEAR_FIELD_FACTOR) if hip_opts is not None else DEFAULT_NEAR_FIELD_FACTOR
    cand = device_near_candidates(dual_to_range, domain, near) if mode == hb.PRECISION_MIXED else None
    srule = None
    if grids_identical:
        srule = singular_rule(dual_to_range.grid, parameters.quadrature.singular, dual_to_range.support, domain.support)

    def regular_only(cfg):
        hb.kbatch_assemble(test, trial, grids_identical, rule, ks, mode, near, cfg, out, parts=hb.KBATCH_REGULAR)

    if config is None:
        key = at.make_key("helmholtz_single_layer", "p0", "p0", rows, cols, nk, precision,
                          torch.cuda.get_device_name(dev))
        config = at.lookup(key)
        if config is None:
            config = at.tune(key, regular_only) if autotune else at.DEFAULT_CONFIG

    if events is None:
        hb.kbatch_assemble(test, trial, grids_identical, rule, ks, mode, near, config, out, near_pairs=cand,
                           singular=srule)
        return out
    events["regular_start"].record()
    hb.kbatch_assemble(test, trial, grids_identical, rule, ks, mode, near, config, out, parts=hb.KBATCH_REGULAR)
    events["regular_end"].record()
    hb.kbatch_assemble(test, trial, grids_identical, rule, ks, mode, near, config, out, near_pairs=cand,
                       singular=srule, parts=hb.KBATCH_NEAR | hb.KBATCH_SINGULAR)
    events["end"].record()
    return out
