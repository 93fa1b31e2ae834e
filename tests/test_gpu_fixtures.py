"""Reference validation tests (test/validation/operators/{boundary,potential}/*.py) run on the HIP path.

Dense boundary matrices and potentials on the reference's 230-triangle sphere.msh are
compared with the reference's golden fixtures at the reference's tolerances
(test/conftest.py:168-174: rtol 1e-10 double, 6e-4 single; Maxwell rtol 1e-10 / atol 1e-14
double, 1e-5 / 1e-7 single, test_maxwell_boundary.py:33-41).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECISIONS = ["double", "single"]


def _boundary(op_name, dom, rng, dual, k, precision, params):
    import bempp_amd.api.operators.boundary as B

    mod, fn = op_name.split(".")
    factory = getattr(getattr(B, mod), fn)
    args = (dom, rng, dual) if k is None else (dom, rng, dual, k)
    return factory(*args, assembler="dense", parameters=params, device_interface="hip", precision=precision)


CASES = [
    # (fixture, operator, domain kind, dual kind, wavenumber)
    ("laplace_single_layer_boundary_p0_p0", "laplace.single_layer", ("DP", 0), ("DP", 0), None),
    ("laplace_single_layer_boundary_p0_dp1", "laplace.single_layer", ("DP", 1), ("DP", 0), None),
    ("laplace_single_layer_boundary_dp1_p0", "laplace.single_layer", ("DP", 0), ("DP", 1), None),
    ("laplace_single_layer_boundary_dp1_dp1", "laplace.single_layer", ("DP", 1), ("DP", 1), None),
    ("laplace_single_layer_boundary_p1_p1", "laplace.single_layer", ("P", 1), ("P", 1), None),
    ("laplace_double_layer_boundary", "laplace.double_layer", ("P", 1), ("P", 1), None),
    ("laplace_adj_double_layer_boundary", "laplace.adjoint_double_layer", ("P", 1), ("P", 1), None),
    ("helmholtz_single_layer_boundary_p0_p0", "helmholtz.single_layer", ("DP", 0), ("DP", 0), 2.5),
    ("helmholtz_single_layer_boundary_p0_dp1", "helmholtz.single_layer", ("DP", 1), ("DP", 0), 2.5),
    ("helmholtz_single_layer_boundary_dp1_p0", "helmholtz.single_layer", ("DP", 0), ("DP", 1), 2.5),
    ("helmholtz_single_layer_boundary_p1_p1", "helmholtz.single_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_double_layer_boundary", "helmholtz.double_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_adj_double_layer_boundary", "helmholtz.adjoint_double_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_complex_single_layer_boundary", "helmholtz.single_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("helmholtz_complex_double_layer_boundary", "helmholtz.double_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("helmholtz_complex_adj_double_layer_boundary", "helmholtz.adjoint_double_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("laplace_hypersingular_boundary", "laplace.hypersingular", ("P", 1), ("P", 1), None),
    ("helmholtz_hypersingular_boundary", "helmholtz.hypersingular", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_complex_hypersingular_boundary", "helmholtz.hypersingular", ("P", 1), ("P", 1), 2.5 + 1j),
]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_boundary_fixture(case, precision, helpers, default_parameters, sphere_grid, hip_available):
    from bempp_amd.api import function_space

    name, op, dk, tk, k = case
    dom = function_space(sphere_grid, *dk)
    dual = function_space(sphere_grid, *tk)
    mat = _boundary(op, dom, dom, dual, k, precision, default_parameters).weak_form().to_dense()
    expected = helpers.load_npy_data(name)
    assert mat.dtype == (np.float32 if precision == "single" else np.float64) or np.iscomplexobj(mat)
    np.testing.assert_allclose(mat, expected, rtol=helpers.default_tolerance(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name,fn,k", [("maxwell_electric_field_boundary", "electric_field", 2.5),
                                       ("maxwell_electric_field_complex_boundary", "electric_field", 2.5 + 1j),
                                       ("maxwell_magnetic_field_boundary", "magnetic_field", 2.5),
                                       ("maxwell_magnetic_field_complex_boundary", "magnetic_field", 2.5 + 1j)])
def test_maxwell_fixture(name, fn, k, precision, helpers, default_parameters, sphere_grid, hip_available):
    import bempp_amd.api.operators.boundary.maxwell as M
    from bempp_amd.api import function_space

    rwg = function_space(sphere_grid, "RWG", 0)
    snc = function_space(sphere_grid, "SNC", 0)
    mat = getattr(M, fn)(rwg, rwg, snc, k, assembler="dense", device_interface="hip", precision=precision,
                         parameters=default_parameters).weak_form().to_dense()
    rtol, atol = (1e-5, 1e-7) if precision == "single" else (1e-10, 1e-14)
    np.testing.assert_allclose(mat, helpers.load_npy_data(name), rtol=rtol, atol=atol)


POTENTIALS = [
    ("laplace_single_layer_potential_p0", "laplace.single_layer", ("DP", 0), None),
    ("laplace_single_layer_potential_p1", "laplace.single_layer", ("P", 1), None),
    ("laplace_double_layer_potential_p1", "laplace.double_layer", ("P", 1), None),
    ("helmholtz_single_layer_potential_p1", "helmholtz.single_layer", ("P", 1), 2.5),
    ("helmholtz_double_layer_potential_p1", "helmholtz.double_layer", ("P", 1), 2.5),
    ("helmholtz_single_layer_potential_complex_p1", "helmholtz.single_layer", ("P", 1), 2.5 + 1j),
    ("helmholtz_double_layer_potential_complex_p1", "helmholtz.double_layer", ("P", 1), 2.5 + 1j),
    ("maxwell_electric_field_potential", "maxwell.electric_field", ("RWG", 0), 2.5),
    ("maxwell_electric_field_potential_complex", "maxwell.electric_field", ("RWG", 0), 2.5 + 1j),
    ("maxwell_magnetic_field_potential", "maxwell.magnetic_field", ("RWG", 0), 2.5),
    ("maxwell_magnetic_field_potential_complex", "maxwell.magnetic_field", ("RWG", 0), 2.5 + 1j),
]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", POTENTIALS, ids=[c[0] for c in POTENTIALS])
def test_potential_fixture(case, precision, helpers, default_parameters, sphere_grid, hip_available):
    import bempp_amd.api.operators.potential as P
    from bempp_amd.api import GridFunction, function_space

    name, op, sk, k = case
    space = function_space(sphere_grid, *sk)
    data = helpers.load_npz_data(name)
    mod, fn = op.split(".")
    factory = getattr(getattr(P, mod), fn)
    args = (space, data["points"]) if k is None else (space, data["points"], k)
    pot = factory(*args, parameters=default_parameters, device_interface="hip", precision=precision)
    actual = pot.evaluate(GridFunction(space, coefficients=data["vec"]))
    np.testing.assert_allclose(actual, data["result"], rtol=helpers.default_tolerance(precision))


FAR_FIELDS = [
    ("helmholtz_single_layer_far_field_p1", "helmholtz.single_layer", ("P", 1)),
    ("helmholtz_double_layer_far_field_p1", "helmholtz.double_layer", ("P", 1)),
    ("maxwell_electric_far_field", "maxwell.electric_field", ("RWG", 0)),
    ("maxwell_magnetic_far_field", "maxwell.magnetic_field", ("RWG", 0)),
]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("case", FAR_FIELDS, ids=[c[0] for c in FAR_FIELDS])
def test_far_field_fixture(case, precision, helpers, default_parameters, sphere_grid, hip_available):
    """test/validation/operators/far_field/test_{helmholtz,maxwell}_far_field.py (k = 2.5)."""
    import bempp_amd.api.operators.far_field as FF
    from bempp_amd.api import GridFunction, function_space

    name, op, sk = case
    space = function_space(sphere_grid, *sk)
    data = helpers.load_npz_data(name)
    mod, fn = op.split(".")
    ff = getattr(getattr(FF, mod), fn)(space, data["points"], 2.5, parameters=default_parameters,
                                       device_interface="hip", precision=precision)
    actual = ff.evaluate(GridFunction(space, coefficients=data["vec"]))
    np.testing.assert_allclose(actual, data["result"], rtol=helpers.default_tolerance(precision))


@pytest.mark.parametrize("case", FAR_FIELDS, ids=[c[0] for c in FAR_FIELDS])
def test_far_field_complex_coefficients(case, helpers, default_parameters, sphere_grid, hip_available):
    """Linearity in the density (test_helmholtz_far_field.py:132-168, test_maxwell_far_field.py:134-170)."""
    import bempp_amd.api.operators.far_field as FF
    from bempp_amd.api import GridFunction, function_space

    name, op, sk = case
    space = function_space(sphere_grid, *sk)
    rng = np.random.default_rng(9)
    c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
    pts = rng.standard_normal((3, 50))
    pts /= np.linalg.norm(pts, axis=0)
    mod, fn = op.split(".")
    ff = getattr(getattr(FF, mod), fn)(space, pts, 2.5, parameters=default_parameters)
    whole = ff.evaluate(GridFunction(space, coefficients=c))
    parts = ff.evaluate(GridFunction(space, coefficients=c.real)) + 1j * ff.evaluate(
        GridFunction(space, coefficients=c.imag))
    np.testing.assert_allclose(whole, parts, rtol=1e-12, atol=1e-14 * np.abs(whole).max())


@pytest.mark.parametrize("precision", PRECISIONS)
def test_laplace_single_element_p0(precision, default_parameters, hip_available):
    """Closed-form value 0.0798 on the unit right triangle (test_laplace_boundary.py:196-226)."""
    from bempp_amd.api import Grid, function_space
    from bempp_amd.api.operators.boundary.laplace import single_layer

    grid = Grid(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64).T, np.array([[0, 1, 2]]).T)
    space = function_space(grid, "DP", 0)
    value = single_layer(space, space, space, assembler="dense", parameters=default_parameters,
                         device_interface="hip", precision=precision).weak_form().A[0, 0]
    assert np.isclose(value, 0.0798, rtol=1e-3)
