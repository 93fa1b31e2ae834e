"""The FP64 oracle against the reference's golden fixtures and against closed forms.

The oracle (oracle/bem_oracle.c) restates the reference's Numba algorithm; these tests pin
it before it is trusted as the checker of the HIP library:
* every dense boundary fixture on sphere.msh at the reference's tolerances
  (test/validation/operators/boundary/*.py, test/conftest.py:168-174);
* every potential fixture (test/validation/operators/potential/*.py) and the far-field fixtures
  of real wavenumber (test/validation/operators/far_field/*.py);
* the closed-form 0.0798 on the unit right triangle (test_laplace_boundary.py:196-226);
* the Laplace single layer on the unit sphere has eigenvalue 1 for constants (V 1 = area
  up to discretisation), the Helmholtz one i k j0(k) h0(k) (spherical Bessel/Hankel),
  which pins the regular + singular split on a mesh no fixture covers;
* the Galerkin single layer of DP0 x DP0 is symmetric for regular pairs.
"""

import numpy as np
import pytest

from oracle import oracle as O

BOUNDARY = [
    ("laplace_single_layer_boundary_p0_p0", "laplace_single_layer", ("DP", 0), ("DP", 0), 0.0),
    ("laplace_single_layer_boundary_p0_dp1", "laplace_single_layer", ("DP", 1), ("DP", 0), 0.0),
    ("laplace_single_layer_boundary_dp1_p0", "laplace_single_layer", ("DP", 0), ("DP", 1), 0.0),
    ("laplace_single_layer_boundary_dp1_dp1", "laplace_single_layer", ("DP", 1), ("DP", 1), 0.0),
    ("laplace_single_layer_boundary_p1_p1", "laplace_single_layer", ("P", 1), ("P", 1), 0.0),
    ("laplace_double_layer_boundary", "laplace_double_layer", ("P", 1), ("P", 1), 0.0),
    ("laplace_adj_double_layer_boundary", "laplace_adjoint_double_layer", ("P", 1), ("P", 1), 0.0),
    ("helmholtz_single_layer_boundary_p0_p0", "helmholtz_single_layer", ("DP", 0), ("DP", 0), 2.5),
    ("helmholtz_single_layer_boundary_p0_dp1", "helmholtz_single_layer", ("DP", 1), ("DP", 0), 2.5),
    ("helmholtz_single_layer_boundary_dp1_p0", "helmholtz_single_layer", ("DP", 0), ("DP", 1), 2.5),
    ("helmholtz_single_layer_boundary_p1_p1", "helmholtz_single_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_double_layer_boundary", "helmholtz_double_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_adj_double_layer_boundary", "helmholtz_adjoint_double_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_complex_single_layer_boundary", "helmholtz_single_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("helmholtz_complex_double_layer_boundary", "helmholtz_double_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("helmholtz_complex_adj_double_layer_boundary", "helmholtz_adjoint_double_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("maxwell_electric_field_boundary", "maxwell_electric_field", ("RWG", 0), ("SNC", 0), 2.5),
    ("maxwell_electric_field_complex_boundary", "maxwell_electric_field", ("RWG", 0), ("SNC", 0), 2.5 + 1j),
    ("laplace_hypersingular_boundary", "laplace_hypersingular", ("P", 1), ("P", 1), 0.0),
    ("helmholtz_hypersingular_boundary", "helmholtz_hypersingular", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_complex_hypersingular_boundary", "helmholtz_hypersingular", ("P", 1), ("P", 1), 2.5 + 1j),
    ("maxwell_magnetic_field_boundary", "maxwell_magnetic_field", ("RWG", 0), ("SNC", 0), 2.5),
    ("maxwell_magnetic_field_complex_boundary", "maxwell_magnetic_field", ("RWG", 0), ("SNC", 0), 2.5 + 1j),
]


@pytest.mark.parametrize("case", BOUNDARY, ids=[c[0] for c in BOUNDARY])
def test_oracle_boundary_fixtures(case, sphere_grid, helpers):
    from bempp_amd.api import function_space

    name, op, dk, tk, k = case
    dom = function_space(sphere_grid, *dk)
    dual = function_space(sphere_grid, *tk)
    A = O.assemble(op, sphere_grid.vertices, sphere_grid.elements, dom, dual, k)
    atol = 1e-14 if op.startswith("maxwell") else 0.0
    np.testing.assert_allclose(A, helpers.load_npy_data(name), rtol=1e-10, atol=atol)


POTENTIALS = [
    ("laplace_single_layer_potential_p0", "laplace_single_layer", ("DP", 0), 0.0),
    ("laplace_single_layer_potential_p1", "laplace_single_layer", ("P", 1), 0.0),
    ("laplace_double_layer_potential_p1", "laplace_double_layer", ("P", 1), 0.0),
    ("helmholtz_single_layer_potential_p1", "helmholtz_single_layer", ("P", 1), 2.5),
    ("helmholtz_double_layer_potential_p1", "helmholtz_double_layer", ("P", 1), 2.5),
    ("helmholtz_single_layer_potential_complex_p1", "helmholtz_single_layer", ("P", 1), 2.5 + 1j),
    ("helmholtz_double_layer_potential_complex_p1", "helmholtz_double_layer", ("P", 1), 2.5 + 1j),
    ("maxwell_electric_field_potential", "maxwell_electric_field", ("RWG", 0), 2.5),
    ("maxwell_electric_field_potential_complex", "maxwell_electric_field", ("RWG", 0), 2.5 + 1j),
    ("maxwell_magnetic_field_potential", "maxwell_magnetic_field", ("RWG", 0), 2.5),
    ("maxwell_magnetic_field_potential_complex", "maxwell_magnetic_field", ("RWG", 0), 2.5 + 1j),
]


# far fields (test/validation/operators/far_field/*.py); the reference's kernels use Re k only, and
# its tests use the fixtures of real wavenumber k = 2.5
FAR_FIELDS = [
    ("helmholtz_single_layer_far_field_p1", "helmholtz_far_field_single_layer", ("P", 1), 2.5),
    ("helmholtz_double_layer_far_field_p1", "helmholtz_far_field_double_layer", ("P", 1), 2.5),
    ("maxwell_electric_far_field", "maxwell_electric_far_field", ("RWG", 0), 2.5),
    ("maxwell_magnetic_far_field", "maxwell_magnetic_far_field", ("RWG", 0), 2.5),
]


@pytest.mark.parametrize("case", POTENTIALS + FAR_FIELDS, ids=[c[0] for c in POTENTIALS + FAR_FIELDS])
def test_oracle_potential_fixtures(case, sphere_grid, helpers):
    from bempp_amd.api import function_space

    name, op, sk, k = case
    sp = function_space(sphere_grid, *sk)
    d = helpers.load_npz_data(name)
    u = O.potential(op, sphere_grid.vertices, sphere_grid.elements, sp, d["vec"], d["points"], k)
    np.testing.assert_allclose(u, d["result"], rtol=1e-10)


def test_oracle_single_triangle_closed_form():
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64).T
    elems = np.array([[0, 1, 2]]).T
    sp = O.SpaceArrays("p0_discontinuous", [[0]], [[1.0]], [1], [1], 1)
    A = O.assemble("laplace_single_layer", verts, elems, sp, sp)
    assert np.isclose(A[0, 0], 0.0798, rtol=1e-3)


def test_oracle_plants_are_caught(sphere_grid, helpers):
    """A wrong singular order or a dropped singular part must break fixture agreement."""
    from bempp_amd.api import function_space

    sp = function_space(sphere_grid, "DP", 0)
    ref = helpers.load_npy_data("laplace_single_layer_boundary_p0_p0")
    no_sing = O.assemble("laplace_single_layer", sphere_grid.vertices, sphere_grid.elements, sp, sp, parts=1)
    assert np.linalg.norm(no_sing - ref) / np.linalg.norm(ref) > 1e-2
    other_order = O.assemble("laplace_single_layer", sphere_grid.vertices, sphere_grid.elements, sp, sp,
                             singular_order=3)
    assert np.max(np.abs(other_order - ref) / np.abs(ref)) > 1e-8


def test_oracle_sphere_eigenvalues():
    """Unit sphere: V 1 = 1 * (M 1) for Laplace and i k j0(k) h0(k) (M 1) for Helmholtz, to O(h^2)."""
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(3)
    sp = function_space(g, "DP", 0)
    mass = g.volumes
    V = O.assemble("laplace_single_layer", g.vertices, g.elements, sp, sp)
    lam = (V @ np.ones(len(mass))).sum() / mass.sum()
    assert abs(lam - 1.0) < 1.5e-2
    k = 2.0
    Vk = O.assemble("helmholtz_single_layer", g.vertices, g.elements, sp, sp, k)
    j0 = np.sin(k) / k
    h0 = -1j * np.exp(1j * k) / k
    lam_k = (Vk @ np.ones(len(mass))).sum() / mass.sum()
    expected = 1j * k * j0 * h0
    assert abs(lam_k - expected) / abs(expected) < 1.5e-2


def test_oracle_regular_part_symmetry():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(2)
    sp = function_space(g, "DP", 0)
    A = O.assemble("helmholtz_single_layer", g.vertices, g.elements, sp, sp, 3.0, parts=1)
    assert np.max(np.abs(A - A.T)) < 1e-15 * np.max(np.abs(A)) * 10
