"""API behaviour that needs no GPU: descriptors, argument checks and error classes
(test/unit/operators/test_boundary_operators.py:117-147)."""

import numpy as np
import pytest


def test_descriptor_fields(sphere_grid):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary import helmholtz

    sp = function_space(sphere_grid, "P", 1)
    op = helmholtz.single_layer(sp, sp, sp, 2.5 + 1j)
    d = op.descriptor
    assert d.options == [2.5, 1.0] and d.kernel_type == "helmholtz_single_layer"
    assert d.assembly_type == "default_scalar" and d.is_complex and d.precision == "double"


def test_efie_space_checks(sphere_grid):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary.maxwell import electric_field

    rwg = function_space(sphere_grid, "RWG", 0)
    snc = function_space(sphere_grid, "SNC", 0)
    p1 = function_space(sphere_grid, "P", 1)
    with pytest.raises(ValueError):
        electric_field(p1, rwg, snc, 1.0)
    with pytest.raises(ValueError):
        electric_field(rwg, rwg, rwg, 1.0)


def test_unknown_interface_and_assembler(sphere_grid):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary import laplace

    sp = function_space(sphere_grid, "DP", 0)
    with pytest.raises(ValueError):
        laplace.single_layer(sp, sp, sp, device_interface="numba").weak_form()
    with pytest.raises(ValueError):
        laplace.single_layer(sp, sp, sp, assembler="fmm")
    with pytest.raises(ValueError):
        laplace.single_layer(sp, sp, sp, precision="half")


def test_modified_helmholtz_rejected(sphere_grid):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary import helmholtz

    sp = function_space(sphere_grid, "DP", 0)
    with pytest.raises(ValueError):
        helmholtz.single_layer(sp, sp, sp, 1j)


def test_unknown_space(sphere_grid):
    from bempp_amd.api import function_space

    with pytest.raises(ValueError):
        function_space(sphere_grid, "DUAL", 0)


def test_singular_pair_arrays(sphere_grid):
    """Pair offsets point at the remapped rule blocks (singular_assembler.py:322-448)."""
    from bempp_amd.api import function_space
    from bempp_amd.core.singular_assembler import singular_rule

    sp = function_space(sphere_grid, "DP", 0)
    r = singular_rule(sphere_grid, 4, sp.support, sp.support, on_device=False)
    n = 6 * 4**4, 5 * 4**4, 2 * 4**4
    assert r.counts == {"coincident": 230, "edge_adjacent": 690, "vertex_adjacent": r.counts["vertex_adjacent"]}
    p = r.pairs
    assert np.all(p[:230, 5] == n[0]) and np.all(p[230:920, 5] == n[1]) and np.all(p[920:, 5] == n[2])
    assert np.all(p[:, 2] + p[:, 5] <= r.test_points.shape[1])
    assert np.all(p[:, 4] + p[:, 5] <= len(r.weights))
    # the remapped test points of an edge pair lie on the shared edge when the Duffy parameter hits it
    el = sphere_grid.elements
    ea = sphere_grid.edge_adjacency[:, 0]
    assert el[ea[2], ea[0]] == el[ea[4], ea[1]]
