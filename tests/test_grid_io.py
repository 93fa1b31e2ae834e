"""gmsh MSH 2.2 export (bempp_cl/api/grid/io.py:39-160) and import round trips, binary and ASCII."""

import os

import numpy as np
import pytest


@pytest.fixture
def grid(helpers):
    from bempp_amd.api.grid.grid import Grid

    g = helpers.load_grid("sphere")
    upper = np.where(g.centroids[:, 2] > 0.0, 7, 3).astype(np.uint32)  # two domains, nonzero tags
    return Grid(g.vertices, g.elements, domain_indices=upper)


@pytest.mark.parametrize("binary", [True, False])
def test_grid_round_trip(grid, tmp_path, binary):
    import bempp_amd.api as bempp

    fn = str(tmp_path / "grid.msh")
    bempp.export(fn, grid=grid, write_binary=binary)
    head = open(fn, "rb").read(40)
    assert head.startswith(b"$MeshFormat\n2.2 " + (b"1" if binary else b"0") + b" 8\n")
    back = bempp.import_grid(fn)
    np.testing.assert_array_equal(back.vertices, grid.vertices)
    np.testing.assert_array_equal(back.elements, grid.elements)
    np.testing.assert_array_equal(back.domain_indices, grid.domain_indices)


def test_reference_fixture_mesh_survives_a_round_trip(helpers, tmp_path):
    """The reference's binary sphere.msh, written again and read back: same vertices and triangles.
    Its physical tags are all 0, so the reader takes the geometrical tags (1 for the one domain),
    the convention of the reference's reader (bempp_cl/api/grid/io.py:27-34)."""
    import bempp_amd.api as bempp

    g = helpers.load_grid("sphere")
    fn = str(tmp_path / "sphere.msh")
    bempp.export(fn, grid=g)
    back = bempp.import_grid(fn)
    np.testing.assert_array_equal(back.vertices, g.vertices)
    np.testing.assert_array_equal(back.elements, g.elements)
    assert set(back.domain_indices.tolist()) == {1}


@pytest.mark.parametrize("binary", [True, False])
def test_grid_function_data(grid, tmp_path, binary):
    import bempp_amd.api as bempp
    from bempp_amd.api.grid.io import read_data_fields

    p1 = bempp.function_space(grid, "P", 1)
    dp0 = bempp.function_space(grid, "DP", 0)
    rng = np.random.default_rng(3)
    f_real = bempp.GridFunction(p1, coefficients=rng.standard_normal(p1.global_dof_count))
    c = rng.standard_normal(dp0.global_dof_count) + 1j * rng.standard_normal(dp0.global_dof_count)
    f_cplx = bempp.GridFunction(dp0, coefficients=c)

    fn = str(tmp_path / "node.msh")
    bempp.export(fn, grid_function=f_real, data_type="node", write_binary=binary)
    fields = read_data_fields(fn)
    np.testing.assert_allclose(fields[("NodeData", "data")][:, 0], f_real.evaluate_on_vertices()[0], rtol=0, atol=0)

    fn = str(tmp_path / "elem.msh")
    bempp.export(fn, grid_function=f_cplx, write_binary=binary)  # default: element data
    fields = read_data_fields(fn)
    centres = f_cplx.evaluate_on_element_centers()[0]
    np.testing.assert_array_equal(fields[("ElementData", "real")][:, 0], centres.real)
    np.testing.assert_array_equal(fields[("ElementData", "imag")][:, 0], centres.imag)
    back = bempp.import_grid(fn)  # the data sections do not disturb the mesh
    np.testing.assert_array_equal(back.elements, grid.elements)

    fn = str(tmp_path / "abs.msh")
    bempp.export(fn, grid_function=f_cplx, transformation="abs", write_binary=binary)
    np.testing.assert_allclose(read_data_fields(fn)[("ElementData", "data")][:, 0], np.abs(centres), rtol=1e-15)


def test_export_argument_errors(grid, tmp_path):
    import bempp_amd.api as bempp

    p1 = bempp.function_space(grid, "P", 1)
    f = bempp.GridFunction(p1, coefficients=np.ones(p1.global_dof_count))
    with pytest.raises(ValueError):
        bempp.export(str(tmp_path / "a.msh"))
    with pytest.raises(ValueError):
        bempp.export(str(tmp_path / "a.msh"), grid=grid, grid_function=f)
    with pytest.raises(ValueError):
        bempp.export(str(tmp_path / "a.msh"), grid_function=f, data_type="vertex")
    with pytest.raises(ValueError):
        bempp.export(str(tmp_path / "a.vtk"), grid=grid)
