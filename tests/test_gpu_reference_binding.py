"""The reference's backend interface on the HIP library (bempp_amd.core.hip_assemblers).

The three backend functions are driven the way the reference's callers drive them, with plain
stand-in objects that carry only the attribute names of the reference's objects:
* assemble_dense (bempp_cl/core/dense_assembler.py:49-100): np.zeros result, dense_assembler
  adds the regular part, singular_assembler fills the flat per-pair values on the localised
  spaces (bempp_cl/core/singular_assembler.py:69-140) and the caller scatters them with
  np.add.at, times the local multipliers;
* DensePotentialAssembler (bempp_cl/core/dense_potential_assembler.py:4-40): the evaluator gets
  map_to_full_grid @ x and its flat result is reshaped (kdim, -1) in Fortran order.
The stand-in grid's data("double") is computed here with numpy as the reference's grid does
(bempp_cl/api/grid/grid.py:552-590). The results must reproduce the reference's fixtures on
sphere.msh at rtol 1e-10.
"""

from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class StandInGrid(object):
    """A grid object exposing only data(precision) (identity equality, like distinct grids)."""

    def __init__(self, vertices, elements):
        v = np.asarray(vertices, dtype=np.float64)
        el = np.asarray(elements, dtype=np.uint32)
        p = v.T[el.T]  # (F, 3 vertices, 3)
        e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
        cross = np.cross(e1, e2)
        norms = np.linalg.norm(cross, axis=1)
        jac = np.stack([e1, e2], axis=2)  # (F, 3, 2)
        jtj = np.einsum("fia,fib->fab", jac, jac)
        self._data = SimpleNamespace(
            vertices=v, elements=el, jacobians=jac, normals=cross / norms[:, None],
            integration_elements=np.sqrt(np.linalg.det(jtj)),
            diameters=np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1) * np.linalg.norm(e1 - e2, axis=1) / norms,
            centroids=p.mean(axis=1))

    def data(self, precision="double"):
        assert precision == "double"
        return self._data


def stand_in_space(space, grid):
    """Plain object with the reference space attributes the backend may read."""

    def plain(s):
        return SimpleNamespace(
            grid=grid, local2global=np.array(s.local2global), local_multipliers=np.array(s.local_multipliers),
            normal_multipliers=np.array(s.normal_multipliers), support_elements=np.array(s.support_elements),
            number_of_shape_functions=s.number_of_shape_functions, global_dof_count=s.global_dof_count,
            shapeset=SimpleNamespace(identifier=s.shapeset.identifier))

    out = plain(space)
    out.localised_space = plain(space.localised_space)
    return out


def stand_in_descriptor(descriptor):
    return SimpleNamespace(kernel_type=descriptor.kernel_type, assembly_type=descriptor.assembly_type,
                           options=list(descriptor.options), precision=descriptor.precision,
                           is_complex=descriptor.is_complex, kernel_dimension=descriptor.kernel_dimension)


def reference_assemble_dense(hip, descriptor, dom, dual, params, grid, singular_arrays):
    """The steps of the reference's assemble_dense with the backend functions of `hip`."""
    dtype = {("single", False): np.float32, ("single", True): np.complex64,
             ("double", False): np.float64, ("double", True): np.complex128}[(descriptor.precision,
                                                                               descriptor.is_complex)]
    result = np.zeros((dual.global_dof_count, dom.global_dof_count), dtype=dtype)
    hip.dense_assembler("hip", descriptor, dom, dual, params, result)
    tp, sp, w, te, se, toff, soff, woff, npts = singular_arrays
    nt, ns = dual.localised_space.number_of_shape_functions, dom.localised_space.number_of_shape_functions
    values = np.zeros(nt * ns * len(te), dtype=dtype)
    hip.singular_assembler("hip", descriptor, grid, dom.localised_space, dual.localised_space, tp, sp, w, te, se,
                           toff, soff, woff, npts, list(descriptor.options), values)
    i_ind = np.tile(np.repeat(np.arange(nt), ns), len(te)) + np.repeat(te * nt, nt * ns)
    j_ind = np.tile(np.tile(np.arange(ns), nt), len(te)) + np.repeat(se * ns, nt * ns)
    rows = dual.local2global.ravel()[i_ind]
    cols = dom.local2global.ravel()[j_ind]
    vals = values * dom.local_multipliers.ravel()[j_ind] * dual.local_multipliers.ravel()[i_ind]
    np.add.at(result, (rows, cols), vals)
    return result


def singular_arrays(sphere_grid, dom, dual, order):
    """The nine arrays of the reference's _SingularQuadratureRuleInterfaceGalerkin.get_arrays()."""
    from bempp_amd.core.singular_assembler import singular_rule

    r = singular_rule(sphere_grid, order, dual.support, dom.support, on_device=False)
    p = r.pairs.astype(np.int64)
    return (r.test_points, r.trial_points, r.weights, p[:, 0], p[:, 1], p[:, 2], p[:, 3], p[:, 4], p[:, 5])


BOUNDARY = [
    ("laplace_single_layer_boundary_p0_p0", "laplace.single_layer", ("DP", 0), ("DP", 0), None),
    ("laplace_single_layer_boundary_p0_dp1", "laplace.single_layer", ("DP", 1), ("DP", 0), None),
    ("laplace_single_layer_boundary_p1_p1", "laplace.single_layer", ("P", 1), ("P", 1), None),
    ("laplace_double_layer_boundary", "laplace.double_layer", ("P", 1), ("P", 1), None),
    ("helmholtz_single_layer_boundary_p1_p1", "helmholtz.single_layer", ("P", 1), ("P", 1), 2.5),
    ("helmholtz_complex_adj_double_layer_boundary", "helmholtz.adjoint_double_layer", ("P", 1), ("P", 1), 2.5 + 1j),
    ("helmholtz_hypersingular_boundary", "helmholtz.hypersingular", ("P", 1), ("P", 1), 2.5),
    ("maxwell_electric_field_boundary", "maxwell.electric_field", ("RWG", 0), ("SNC", 0), 2.5),
    ("maxwell_magnetic_field_complex_boundary", "maxwell.magnetic_field", ("RWG", 0), ("SNC", 0), 2.5 + 1j),
]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", BOUNDARY, ids=[c[0] for c in BOUNDARY])
def test_boundary_through_reference_interface(case, precision, helpers, default_parameters, sphere_grid,
                                              hip_available):
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from bempp_amd.core import hip_assemblers

    name, op, dk, tk, k = case
    dom = function_space(sphere_grid, *dk)
    dual = function_space(sphere_grid, *tk)
    mod, fn = op.split(".")
    factory = getattr(getattr(B, mod), fn)
    args = (dom, dom, dual) if k is None else (dom, dom, dual, k)
    descriptor = stand_in_descriptor(factory(*args, precision=precision).descriptor)
    grid = StandInGrid(sphere_grid.vertices, sphere_grid.elements)
    sdom, sdual = stand_in_space(dom, grid), stand_in_space(dual, grid)
    params = SimpleNamespace(quadrature=SimpleNamespace(regular=4, singular=4))
    mat = reference_assemble_dense(hip_assemblers, descriptor, sdom, sdual, params, grid,
                                   singular_arrays(sphere_grid, dom, dual, 4))
    expected = helpers.load_npy_data(name)
    if mod == "maxwell":
        rtol, atol = (1e-5, 1e-7) if precision == "single" else (1e-10, 1e-14)
        np.testing.assert_allclose(mat, expected, rtol=rtol, atol=atol)
    else:
        np.testing.assert_allclose(mat, expected, rtol=helpers.default_tolerance(precision))


POTENTIALS = [
    ("laplace_single_layer_potential_p0", "laplace.single_layer", ("DP", 0), None),
    ("helmholtz_double_layer_potential_complex_p1", "helmholtz.double_layer", ("P", 1), 2.5 + 1j),
    ("maxwell_electric_field_potential", "maxwell.electric_field", ("RWG", 0), 2.5),
]


@pytest.mark.parametrize("case", POTENTIALS, ids=[c[0] for c in POTENTIALS])
def test_potential_through_reference_interface(case, helpers, sphere_grid, hip_available):
    import bempp_amd.api.operators.potential as P
    from bempp_amd.api import function_space
    from bempp_amd.core import hip_assemblers

    name, op, sk, k = case
    space = function_space(sphere_grid, *sk)
    data = helpers.load_npz_data(name)
    mod, fn = op.split(".")
    factory = getattr(getattr(P, mod), fn)
    args = (space, data["points"]) if k is None else (space, data["points"], k)
    descriptor = stand_in_descriptor(factory(*args, precision="double").descriptor)
    grid = StandInGrid(sphere_grid.vertices, sphere_grid.elements)
    sspace = stand_in_space(space, grid)
    params = SimpleNamespace(quadrature=SimpleNamespace(regular=4, singular=4))
    evaluator = hip_assemblers.potential_assembler("hip", sspace.localised_space, descriptor, data["points"], params)
    x = space.map_to_full_grid @ np.asarray(data["vec"])
    actual = evaluator(x).reshape([descriptor.kernel_dimension, -1], order="F")
    np.testing.assert_allclose(actual, data["result"], rtol=1e-10)


def test_reference_interface_rejects_other_interfaces(sphere_grid, hip_available):
    from bempp_amd.core import hip_assemblers

    with pytest.raises(ValueError):
        hip_assemblers.dense_assembler("numba", None, None, None, None, None)
