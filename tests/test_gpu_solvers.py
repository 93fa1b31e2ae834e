"""Solves on operators assembled by the HIP path.

* The reference's mixed Dirichlet / Neumann Laplace problem on cube_grid.msh
  (test/validation/computations/test_mixed_neumann_dirichlet_.py): a 2 x 2 blocked system of
  single, double, adjoint double layer and hypersingular operators with the L2 identity,
  right-hand sides projected from callables, GMRES, then the solution on the element centres
  against the reference's golden values (mixed_dirichlet_neumann_sol.npz) at the reference's
  tolerances (rtol 1e-10 double, 6e-4 single).
* Device GMRES / CG / LU against scipy / numpy on the same assembled matrices (host copies).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["double", "single"])
def test_laplace_mixed_neumann_dirichlet(precision, helpers, default_parameters, hip_available):
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import laplace, sparse

    grid = helpers.load_grid("cube_grid")
    dirichlet_segments = [1, 3]
    neumann_segments = [2, 4, 5, 6]
    fs = bempp.function_space
    global_neumann_space = fs(grid, "DP", 0)
    global_dirichlet_space = fs(grid, "P", 1)
    neumann_space_dirichlet_segment = fs(grid, "DP", 0, segments=dirichlet_segments)
    neumann_space_neumann_segment = fs(grid, "DP", 0, segments=neumann_segments)
    dirichlet_space_dirichlet_segment = fs(grid, "P", 1, segments=dirichlet_segments, include_boundary_dofs=True,
                                           truncate_at_segment_edge=False)
    dirichlet_space_neumann_segment = fs(grid, "P", 1, segments=neumann_segments)
    dual_dirichlet_space = fs(grid, "P", 1, segments=dirichlet_segments, include_boundary_dofs=True)
    kw = dict(parameters=default_parameters, precision=precision)
    slp_DD = laplace.single_layer(neumann_space_dirichlet_segment, dirichlet_space_dirichlet_segment,
                                  neumann_space_dirichlet_segment, **kw)
    dlp_DN = laplace.double_layer(dirichlet_space_neumann_segment, dirichlet_space_dirichlet_segment,
                                  neumann_space_dirichlet_segment, **kw)
    adlp_ND = laplace.adjoint_double_layer(neumann_space_dirichlet_segment, neumann_space_neumann_segment,
                                           dirichlet_space_neumann_segment, **kw)
    hyp_NN = laplace.hypersingular(dirichlet_space_neumann_segment, neumann_space_neumann_segment,
                                   dirichlet_space_neumann_segment, **kw)
    slp_DN = laplace.single_layer(neumann_space_neumann_segment, dirichlet_space_dirichlet_segment,
                                  neumann_space_dirichlet_segment, **kw)
    dlp_DD = laplace.double_layer(dirichlet_space_dirichlet_segment, dirichlet_space_dirichlet_segment,
                                  neumann_space_dirichlet_segment, **kw)
    id_DD = sparse.identity(dirichlet_space_dirichlet_segment, dirichlet_space_dirichlet_segment,
                            neumann_space_dirichlet_segment)
    adlp_NN = laplace.adjoint_double_layer(neumann_space_neumann_segment, neumann_space_neumann_segment,
                                           dirichlet_space_neumann_segment, **kw)
    id_NN = sparse.identity(neumann_space_neumann_segment, neumann_space_neumann_segment,
                            dirichlet_space_neumann_segment)
    hyp_ND = laplace.hypersingular(dirichlet_space_dirichlet_segment, neumann_space_neumann_segment,
                                   dirichlet_space_neumann_segment, **kw)

    blocked = bempp.BlockedOperator(2, 2)
    blocked[0, 0] = slp_DD
    blocked[0, 1] = -dlp_DN
    blocked[1, 0] = adlp_ND
    blocked[1, 1] = hyp_NN

    @bempp.real_callable
    def dirichlet_data(x, n, domain_index, res):
        res[0] = 1

    @bempp.real_callable
    def neumann_data(x, n, domain_index, res):
        res[0] = 1

    dirichlet_grid_fun = bempp.GridFunction(dirichlet_space_dirichlet_segment, fun=dirichlet_data,
                                            dual_space=dual_dirichlet_space)
    neumann_grid_fun = bempp.GridFunction(neumann_space_neumann_segment, fun=neumann_data,
                                          dual_space=dirichlet_space_neumann_segment)
    rhs_fun1 = (0.5 * id_DD + dlp_DD) * dirichlet_grid_fun - slp_DN * neumann_grid_fun
    rhs_fun2 = -hyp_ND * dirichlet_grid_fun + (0.5 * id_NN - adlp_NN) * neumann_grid_fun

    (neumann_solution, dirichlet_solution), info, its = bempp.linalg.gmres(
        blocked, [rhs_fun1, rhs_fun2], return_iteration_count=True)
    assert info == 0

    embed = sparse.identity
    dirichlet = (embed(dirichlet_space_dirichlet_segment, global_dirichlet_space, global_dirichlet_space)
                 * dirichlet_grid_fun
                 + embed(dirichlet_space_neumann_segment, global_dirichlet_space, global_dirichlet_space)
                 * dirichlet_solution)
    neumann = (embed(neumann_space_neumann_segment, global_neumann_space, global_neumann_space) * neumann_grid_fun
               + embed(neumann_space_dirichlet_segment, global_neumann_space, global_neumann_space) * neumann_solution)
    data = helpers.load_npz_data("mixed_dirichlet_neumann_sol")
    actual_dirichlet = dirichlet.evaluate_on_element_centers()
    actual_neumann = neumann.evaluate_on_element_centers()
    tol = helpers.default_tolerance(precision)
    print(f"{precision}: {its} GMRES iterations; max rel err dirichlet "
          f"{np.max(np.abs(actual_dirichlet - data['dirichlet_vals']) / np.abs(data['dirichlet_vals'])):.2e}, "
          f"neumann {np.max(np.abs(actual_neumann - data['neumann_vals']) / np.abs(data['neumann_vals'])):.2e}",
          flush=True)
    np.testing.assert_allclose(actual_dirichlet, data["dirichlet_vals"], rtol=tol)
    np.testing.assert_allclose(actual_neumann, data["neumann_vals"], rtol=tol)


@pytest.fixture(scope="module")
def helmholtz_system():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary import helmholtz

    grid = icosphere(3)
    p1 = function_space(grid, "P", 1)
    return grid, p1, helmholtz


@pytest.mark.parametrize("precision,restart", [("double", None), ("single", 10)])
def test_gmres_vs_scipy_on_assembled_matrix(precision, restart, helmholtz_system, default_parameters,
                                            hip_available):
    import scipy.sparse.linalg as sla

    import bempp_amd.api as bempp

    grid, p1, helmholtz = helmholtz_system
    op = helmholtz.single_layer(p1, p1, p1, 3.0, parameters=default_parameters, precision=precision)

    @bempp.complex_callable
    def wave(x, n, domain_index, res):
        res[0] = np.exp(3.0j * x[0])

    rhs = bempp.GridFunction(p1, fun=wave)
    sol, info, res, its = bempp.linalg.gmres(op, rhs, tol=1e-8, restart=restart, return_residuals=True,
                                             return_iteration_count=True)
    A = op.weak_form().to_dense()
    b = rhs.projections()
    ref_res = []
    x_ref, info_ref = sla.gmres(A, b, rtol=1e-8, restart=restart, callback=ref_res.append, callback_type="legacy")
    print(f"{precision}: {its} iterations (scipy {len(ref_res)})", flush=True)
    assert info == info_ref == 0
    assert its == len(ref_res) == len(res)
    np.testing.assert_allclose(res, ref_res, rtol=1e-5, atol=1e-12)
    x = sol.coefficients
    # the same Krylov iterate: differences come from the products' rounding (FP64 accumulation
    # on the device, the matrix's dtype in numpy)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < (1e-9 if precision == "double" else 1e-5)
    assert np.linalg.norm(A.astype(np.complex128) @ x - b) <= 1e-8 * np.linalg.norm(b) * 1.01


def test_cg_and_lu_on_laplace_single_layer(default_parameters, hip_available):
    import scipy.sparse.linalg as sla

    import bempp_amd.api as bempp
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary import laplace

    grid = icosphere(3)
    dp0 = function_space(grid, "DP", 0)
    op = laplace.single_layer(dp0, dp0, dp0, parameters=default_parameters)

    @bempp.real_callable
    def data(x, n, domain_index, res):
        res[0] = x[2] + 0.5 * x[0] * x[1]

    rhs = bempp.GridFunction(dp0, fun=data)
    sol, info, its = bempp.linalg.cg(op, rhs, tol=1e-10, return_iteration_count=True)
    A = op.weak_form().to_dense()
    b = rhs.projections()
    xs = []
    x_ref, info_ref = sla.cg(A, b, rtol=1e-10, callback=xs.append)
    assert info == info_ref == 0 and its == len(xs)
    assert sol.coefficients.dtype == np.float64
    np.testing.assert_allclose(sol.coefficients, x_ref, rtol=1e-8, atol=1e-10 * np.abs(x_ref).max())
    lu_sol = bempp.linalg.lu(op, rhs)
    x_np = np.linalg.solve(A, b)
    np.testing.assert_allclose(lu_sol.coefficients, x_np, rtol=1e-9, atol=1e-12 * np.abs(x_np).max())
    factors = bempp.linalg.compute_lu_factors(op)
    np.testing.assert_allclose(bempp.linalg.lu(op, rhs, lu_factor=factors).coefficients, x_np, rtol=1e-9,
                               atol=1e-12 * np.abs(x_np).max())


@pytest.mark.parametrize("dtype", ["float64", "complex64"])
def test_discrete_operator_products(dtype, default_parameters, hip_available):
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary import helmholtz, laplace, sparse

    grid = icosphere(2)
    p1 = function_space(grid, "P", 1)
    if dtype == "float64":
        op = laplace.double_layer(p1, p1, p1, parameters=default_parameters)
    else:
        op = helmholtz.double_layer(p1, p1, p1, 2.0, parameters=default_parameters, precision="single")
    W = op.weak_form()
    A = W.to_dense()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(A.shape[1]) + 1j * rng.standard_normal(A.shape[1])
    np.testing.assert_allclose(W @ x, A.astype(np.complex128) @ x, rtol=1e-12, atol=1e-14)
    xr = rng.standard_normal(A.shape[1])
    y = W @ xr
    assert y.dtype == np.result_type(A.dtype, xr.dtype)
    S = (0.5 * sparse.identity(p1, p1, p1) + op).weak_form()
    M = sparse.identity(p1, p1, p1).weak_form().to_dense()
    np.testing.assert_allclose(S.to_dense(), (0.5 * M + A).astype(S.dtype), rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("k,n", [(1, 37), (5, 1000), (21, 10242), (40, 200000), (300, 2000)])
def test_gmres_cgs2_entry_point(k, n, hip_available):
    """bem_gmres_cgs2 against the same CGS2 arithmetic in numpy (FP64): the coefficients, both
    norms, the updated vector and the next basis vector (row k of V: w2 / |w2|, or w2 itself when
    the breakdown test |w2| <= eps |w| holds); n past one block, past the 512-block grid cap
    (grid-stride), k past the restart length used elsewhere and past the 256 dot products staged
    per barrier."""
    import torch

    from bempp_amd.core import hip_backend as hb

    rng = np.random.default_rng(k + n)
    V = rng.standard_normal((k + 2, n)) + 1j * rng.standard_normal((k + 2, n))
    V[:k] = np.linalg.qr(V[:k].T)[0].T  # orthonormal rows, as in GMRES
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c1 = V[:k].conj() @ w
    w1 = w - V[:k].T @ c1
    c2 = V[:k].conj() @ w1
    w2 = w1 - V[:k].T @ c2
    dev = hb.device()
    Vt = torch.as_tensor(V, device=dev)
    wt = torch.as_tensor(w, device=dev)
    out = torch.full((2 * k + 2,), 7.0 + 7.0j, dtype=torch.complex128, device=dev)  # overwritten
    hb.gmres_cgs2(Vt, k, wt, out, Vt[k], 1e-14)
    g = out.cpu().numpy()
    scale = np.linalg.norm(w)
    np.testing.assert_allclose(g[:k], c1, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(g[k + 1 : 2 * k + 1], c2, rtol=0, atol=1e-12 * scale)
    assert abs(g[k].real - np.vdot(w, w).real) <= 1e-12 * scale**2 and g[k].imag == 0.0
    assert abs(g[2 * k + 1].real - np.vdot(w2, w2).real) <= 1e-12 * scale**2
    np.testing.assert_allclose(wt.cpu().numpy(), w2, rtol=0, atol=1e-12 * scale)
    Vh = Vt.cpu().numpy()
    np.testing.assert_allclose(Vh[k], w2 / np.linalg.norm(w2), rtol=0, atol=1e-12)
    # the other rows are untouched
    np.testing.assert_array_equal(Vh[:k], V[:k])
    np.testing.assert_array_equal(Vh[k + 1], V[k + 1])
    # breakdown test (eps so large that it holds): row k + 1 receives w_out unscaled
    w3 = wt.clone()
    hb.gmres_cgs2(Vt, k, w3, out, Vt[k + 1], 1e300)
    np.testing.assert_array_equal(Vt[k + 1].cpu().numpy(), w3.cpu().numpy())

