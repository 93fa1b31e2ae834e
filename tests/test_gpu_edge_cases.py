"""Edge cases of the C ABI and the host wrappers on the device: a space without elements, an empty
point set, a coefficient vector of the wrong length."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op,kind,degree,kdim", [("helmholtz_single_layer", "P", 1, 1),
                                                 ("maxwell_electric_field", "RWG", 0, 3)])
@pytest.mark.parametrize("precision_name", ["fp64", "mixed"])
def test_potential_of_empty_space_is_zero(op, kind, degree, kdim, precision_name, hip_available):
    """A space restricted to no elements has no sources: bem_potential_evaluate writes zeros
    (the workspace plan and the launches must not divide by or launch with a zero count)."""
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb

    L = hb.lib()
    dev = hb.device()
    space = bempp.function_space(bempp.shapes.icosphere(1), kind, degree)
    dspace = hb.DeviceSpace(space, elements=np.zeros(0, dtype=np.int64))
    assert dspace.struct.n_elements == 0
    precision = hb.PRECISION_MIXED if precision_name == "mixed" else hb.PRECISION_FP64
    out_dtype = torch.complex64 if precision == hb.PRECISION_MIXED else torch.complex128
    n_points = 300
    pts = torch.from_numpy(np.random.default_rng(0).standard_normal((n_points, 3))).to(dev)
    coeffs = torch.zeros(space.grid.number_of_elements * 3, dtype=torch.complex128, device=dev)
    qp, qw = triangle_rule(4)
    rule = hb.quad_rule(qp, qw)
    opc = hb.OPERATORS[op]
    need = L.bem_potential_workspace_size(opc, ctypes.byref(dspace.struct), ctypes.byref(rule), n_points, precision)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    out = torch.full((kdim, n_points), float("nan"), dtype=out_dtype, device=dev)
    hb.check(L.bem_potential_evaluate(opc, ctypes.byref(dspace.struct), hb.ptr(coeffs), n_points, hb.ptr(pts),
                                      ctypes.byref(rule), 2.0, 0.0, precision, hb.ptr(out),
                                      hb.DTYPES[str(out_dtype).replace("torch.", "")], hb.ptr(ws), need,
                                      hb.stream_handle()))
    torch.cuda.synchronize()
    assert torch.all(out == 0).item()
    # and no points at all: nothing to write, no error, no pointers needed
    hb.check(L.bem_potential_evaluate(opc, ctypes.byref(dspace.struct), hb.ptr(coeffs), 0, None,
                                      ctypes.byref(rule), 2.0, 0.0, precision, None,
                                      hb.DTYPES[str(out_dtype).replace("torch.", "")], hb.ptr(ws), need,
                                      hb.stream_handle()))


def test_potential_rejects_wrong_coefficient_count(hip_available):
    """A coefficient vector of the wrong length is refused on the host, before the device gather
    that maps it to the elements."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.potential import helmholtz as hpot

    space = bempp.function_space(bempp.shapes.icosphere(1), "P", 1)
    pts = np.array([[2.0, 0.0, 0.0], [0.0, 2.0, 0.0]]).T
    pot = hpot.single_layer(space, pts, 1.5)
    impl = pot._assembler._implementation._implementation
    with pytest.raises(ValueError):
        impl(np.ones(space.global_dof_count + 1))
    u = impl(np.ones(space.global_dof_count))
    assert u.shape == (1, 2) and np.all(np.isfinite(u))
