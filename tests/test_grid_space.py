"""Grid geometry/adjacency and function-space dof maps.

Pins: sphere.msh has 117 vertices / 230 triangles / 345 edges; icospheres satisfy
V - E + F = 2, unit-norm vertices, outward normals, and the singular pair counts
coincident = F, edge = 3F, vertex = 120 + 18 (V - 12); every adjacency entry names vertices
that really are shared; P1 and RWG dof maps equal the oracle's independent closed-mesh maps.
"""

import numpy as np
import pytest


def test_sphere_msh(sphere_grid):
    g = sphere_grid
    assert (g.number_of_vertices, g.number_of_elements, g.number_of_edges) == (117, 230, 345)


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_icosphere_topology(level):
    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(level)
    F, V, E = g.number_of_elements, g.number_of_vertices, g.number_of_edges
    assert F == 20 * 4**level and V == 10 * 4**level + 2
    assert V - E + F == 2
    np.testing.assert_allclose(np.linalg.norm(g.vertices, axis=0), 1.0, atol=1e-15)
    assert np.all(np.sum(g.normals * g.centroids, axis=1) > 0)
    assert g.edge_adjacency.shape[1] == 3 * F
    assert g.vertex_adjacency.shape[1] == 120 + 18 * (V - 12)


def test_adjacency_consistency(sphere_grid):
    g = sphere_grid
    el = g.elements
    ea = g.edge_adjacency
    for col in ea.T:
        e0, e1, a0, a1, b0, b1 = col
        assert el[a0, e0] == el[b0, e1] and el[a1, e0] == el[b1, e1]
        assert a0 != a1 and b0 <= b1
    va = g.vertex_adjacency
    for col in va.T:
        e0, e1, a, b = col
        assert el[a, e0] == el[b, e1]
        assert len(set(el[:, e0]) & set(el[:, e1])) == 1


def test_geometry(sphere_grid):
    g = sphere_grid
    v = g.vertices
    el = g.elements
    j = g.jacobians
    np.testing.assert_allclose(j[:, :, 0], (v[:, el[1]] - v[:, el[0]]).T)
    cross = np.cross(j[:, :, 0], j[:, :, 1])
    np.testing.assert_allclose(g.integration_elements, np.linalg.norm(cross, axis=1))
    np.testing.assert_allclose(g.volumes, 0.5 * g.integration_elements)


def test_refine_children():
    from bempp_amd.api.grid.shapes import icosahedron

    g = icosahedron()
    r = g.refine()
    assert r.number_of_elements == 80
    # children of element 0: (v0, m01, m20), (m01, v1, m12), (m12, v2, m20), (m01, m12, m20)
    v0, v1, v2 = g.elements[:, 0]
    assert r.elements[0, 0] == v0 and r.elements[1, 1] == v1 and r.elements[1, 2] == v2
    np.testing.assert_allclose(r.volumes[:4].sum(), g.volumes[0])


def test_p1_rwg_dofmaps_match_independent_maps(sphere_grid):
    from bempp_amd.api import function_space
    from oracle import oracle as O

    g = sphere_grid
    p1 = function_space(g, "P", 1)
    ref = O.p1_closed_dofmap(g.elements, g.number_of_vertices)
    assert np.array_equal(p1.local2global.astype(np.int32), ref.local2global)
    assert p1.global_dof_count == 117
    rwg = function_space(g, "RWG", 0)
    r = O.rwg_closed_dofmap(g.elements)
    assert np.array_equal(rwg.local2global.astype(np.int32), r.local2global)
    assert np.array_equal(rwg.local_multipliers, r.local_multipliers)
    assert rwg.global_dof_count == 345


def test_dp_spaces(sphere_grid):
    from bempp_amd.api import function_space

    dp0 = function_space(sphere_grid, "DP", 0)
    dp1 = function_space(sphere_grid, "DP", 1)
    assert dp0.global_dof_count == 230 and dp1.global_dof_count == 690
    assert dp0.is_localised and dp1.is_localised
    p1 = function_space(sphere_grid, "P", 1)
    assert not p1.is_localised and p1.localised_space.global_dof_count == 690


def test_segments_support(sphere_grid):
    from bempp_amd.api import function_space
    from bempp_amd.api.grid import Grid

    g = sphere_grid
    dom = np.arange(g.number_of_elements) % 2
    g2 = Grid(g.vertices, g.elements, domain_indices=dom)
    sp = function_space(g2, "DP", 0, segments=[1])
    assert sp.global_dof_count == int(dom.sum())
    sw = function_space(g2, "DP", 0, swapped_normals=[1])
    assert np.all(sw.normal_multipliers[dom == 1] == -1) and np.all(sw.normal_multipliers[dom == 0] == 1)
    with pytest.raises(ValueError):
        function_space(g2, "DP", 0, segments=[1], support_elements=[0])
    p1 = function_space(g2, "P", 1, segments=[1])
    assert p1.global_dof_count < g.number_of_vertices


def test_gmsh_roundtrip(tmp_path, sphere_grid):
    from bempp_amd.api.grid import export, import_grid

    fname = str(tmp_path / "s.msh")
    export(fname, grid=sphere_grid)
    g = import_grid(fname)
    np.testing.assert_allclose(g.vertices, sphere_grid.vertices)
    assert np.array_equal(g.elements, sphere_grid.elements)
