"""bem_potential_evaluate_noalloc (library-owned scratch buffer, no allocation) against
bem_potential_evaluate with a caller workspace, through the C ABI.

Both entry points run the same kernels with the same chunk plan and a fixed-order chunk
reduction, so their results must agree bit for bit. Two calls in a row, the second on a
smaller input, check that what the first call left in the shared scratch buffer does not leak
into the second.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _evaluate_both(op, space, k, precision, n_points, seed):
    import torch

    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb

    L = hb.lib()
    dev = hb.device()
    dspace = hb.device_space(space)
    nshape = {"p0_discontinuous": 1, "p1_discontinuous": 3, "rwg0": 3, "snc0": 3}[space.shapeset.identifier]
    rng = np.random.default_rng(seed)
    n_coef = space.grid.number_of_elements * nshape
    coeffs = torch.from_numpy(rng.standard_normal(n_coef) + 1j * rng.standard_normal(n_coef)).to(dev)
    pts = rng.standard_normal((n_points, 3))
    pts = 2.0 * pts / np.linalg.norm(pts, axis=1)[:, None]
    dpts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    qp, qw = triangle_rule(4)
    rule = hb.quad_rule(qp, qw)
    kdim = 3 if op.startswith("maxwell") else 1
    out_dtype = torch.complex64 if precision == hb.PRECISION_MIXED else torch.complex128
    code = hb.DTYPES[str(out_dtype).replace("torch.", "")]
    opc = hb.OPERATORS[op]
    need = L.bem_potential_workspace_size(opc, ctypes.byref(dspace.struct), ctypes.byref(rule), n_points, precision)
    assert 0 < need <= (64 << 20)
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    ref = torch.full((kdim, n_points), float("nan"), dtype=out_dtype, device=dev)
    out = torch.full((kdim, n_points), float("nan"), dtype=out_dtype, device=dev)
    stream = hb.stream_handle()
    hb.check(L.bem_potential_evaluate(opc, ctypes.byref(dspace.struct), hb.ptr(coeffs), n_points, hb.ptr(dpts),
                                      ctypes.byref(rule), k.real, k.imag, precision, hb.ptr(ref), code, hb.ptr(ws),
                                      need, stream))
    hb.check(L.bem_potential_evaluate_noalloc(opc, ctypes.byref(dspace.struct), hb.ptr(coeffs), n_points,
                                              hb.ptr(dpts), ctypes.byref(rule), k.real, k.imag, precision,
                                              hb.ptr(out), code, stream))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


@pytest.mark.parametrize("precision_name", ["fp64", "mixed"])
@pytest.mark.parametrize("op,kind,degree", [("helmholtz_single_layer", "P", 1),
                                            ("helmholtz_double_layer", "DP", 0),
                                            ("maxwell_electric_field", "RWG", 0)])
def test_noalloc_matches_workspace_entry(op, kind, degree, precision_name):
    import bempp_amd.api as bempp
    from bempp_amd.core import hip_backend as hb

    precision = hb.PRECISION_FP64 if precision_name == "fp64" else hb.PRECISION_MIXED
    k = complex(2.5, 0.0) if op != "helmholtz_double_layer" else complex(2.0, 0.5)
    # a larger call, then a smaller one: the scratch buffer keeps the first call's contents
    for level, n_points, seed in [(3, 3000, 1), (2, 700, 2)]:
        space = bempp.function_space(bempp.shapes.icosphere(level), kind, degree)
        ref, out = _evaluate_both(op, space, k, precision, n_points, seed)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(out, ref)


def test_noalloc_rejects_oversize_request():
    """A potential whose workspace exceeds the scratch buffer is refused before any device work."""
    import bempp_amd.api as bempp
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb

    L = hb.lib()
    space = bempp.function_space(bempp.shapes.icosphere(2), "P", 1)
    dspace = hb.device_space(space)
    qp, qw = triangle_rule(4)
    rule = hb.quad_rule(qp, qw)
    n_points = 1 << 23
    need = L.bem_potential_workspace_size(3, ctypes.byref(dspace.struct), ctypes.byref(rule), n_points, 1)
    assert need > (64 << 20)
    dummy = ctypes.c_void_p(256)
    rc = L.bem_potential_evaluate_noalloc(3, ctypes.byref(dspace.struct), dummy, n_points, dummy,
                                          ctypes.byref(rule), 1.0, 0.0, 1, dummy, 2, hb.stream_handle())
    assert rc == 4
    assert b"scratch" in L.bem_last_error()
