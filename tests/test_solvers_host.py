"""Host logic of the solve path: the Krylov restatements, the mass matrix, grid functions.

* GMRES / CG (bempp_amd.api.linalg.iterative_solvers) run on CPU tensors with a CPU product
  and are compared with scipy.sparse.linalg.gmres / cg, which the reference calls
  (bempp_cl/api/linalg/iterative_solvers.py:137-140, 189-190): same iteration count, same
  residual history (the legacy callback values) and the same iterate to rounding.
* The mass matrix (the L2 identity's weak form) against closed forms: DP0 x DP0 is diag(area),
  P1 x P1 sums to the total area and has the element matrix area/12 [2 1 1; 1 2 1; 1 1 2],
  DP0 x P1 has row sums area and c^T M 1 = integral of c.
* Grid functions: projecting a callable, coefficients <-> projections round trips, evaluation.
"""

import numpy as np
import pytest


def _system(n, seed, complex_=True, spd=False):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if complex_ else 0)
    if spd:
        A = A.conj().T @ A + n * np.eye(n)
    else:
        A = A / np.sqrt(n) + 2.0 * np.eye(n)
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0)
    return A, b


@pytest.mark.parametrize("n,restart,tol,complex_", [(60, None, 1e-5, True), (80, 7, 1e-9, True), (50, 5, 1e-8, False),
                                                    (30, 40, 1e-10, True)])
def test_gmres_matches_scipy(n, restart, tol, complex_):
    import scipy.sparse.linalg as sla
    import torch

    from bempp_amd.api.linalg.iterative_solvers import krylov_gmres

    A, b = _system(n, n + (restart or 0), complex_)
    ref_res = []
    x_ref, info_ref = sla.gmres(A, b, rtol=tol, restart=restart, callback=ref_res.append, callback_type="legacy")
    At = torch.as_tensor(A.astype(np.complex128))
    res = []
    x, info = krylov_gmres(lambda v: At @ v, b, torch.device("cpu"), tol, restart, None, res.append)
    assert info == info_ref == 0
    assert len(res) == len(ref_res)
    np.testing.assert_allclose(res, ref_res, rtol=1e-7, atol=1e-14)
    x = x.numpy()
    np.testing.assert_allclose(x, x_ref, rtol=1e-9, atol=1e-12 * np.linalg.norm(x_ref))
    assert np.linalg.norm(A @ x - b) <= tol * np.linalg.norm(b) * (1 + 1e-9)


def test_gmres_maxiter_and_zero_rhs():
    import scipy.sparse.linalg as sla
    import torch

    from bempp_amd.api.linalg.iterative_solvers import krylov_gmres

    A, b = _system(40, 3)
    At = torch.as_tensor(A)
    x, info = krylov_gmres(lambda v: At @ v, np.zeros(40), torch.device("cpu"))
    assert info == 0 and float(torch.linalg.vector_norm(x)) == 0.0
    _, info_ref = sla.gmres(A, b, rtol=1e-14, restart=3, maxiter=2, callback=lambda r: None, callback_type="legacy")
    _, info = krylov_gmres(lambda v: At @ v, b, torch.device("cpu"), 1e-14, 3, 2)
    assert info == info_ref != 0


@pytest.mark.parametrize("complex_", [False, True])
def test_cg_matches_scipy(complex_):
    import scipy.sparse.linalg as sla
    import torch

    from bempp_amd.api.linalg.iterative_solvers import krylov_cg

    A, b = _system(70, 5, complex_, spd=True)
    xs = []
    x_ref, info_ref = sla.cg(A, b, rtol=1e-10, callback=lambda xk: xs.append(xk.copy()))
    At = torch.as_tensor(A.astype(np.complex128))
    mine = []
    x, info = krylov_cg(lambda v: At @ v, b, torch.device("cpu"), 1e-10, None, lambda xk: mine.append(xk.clone()))
    assert info == info_ref == 0 and len(mine) == len(xs)
    np.testing.assert_allclose(x.numpy(), x_ref, rtol=1e-9, atol=1e-12)


def test_lartg_rotation():
    from bempp_amd.api.linalg.iterative_solvers import _lartg

    rng = np.random.default_rng(0)
    for _ in range(20):
        f, g = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        c, s, r = _lartg(f, g)
        assert abs(c * f + s * g - r) < 1e-14 and abs(-np.conj(s) * f + c * g) < 1e-14
        assert abs(c * c + abs(s) ** 2 - 1) < 1e-14
    assert _lartg(2.0, 0.0)[0] == 1.0 and _lartg(0.0, 3.0)[0] == 0.0


@pytest.fixture(scope="module")
def ico():
    from bempp_amd.api.grid.shapes import icosphere

    return icosphere(2)


def test_mass_matrix_closed_forms(ico):
    from bempp_amd.api import function_space
    from bempp_amd.core.sparse_assembler import mass_matrix

    dp0 = function_space(ico, "DP", 0)
    p1 = function_space(ico, "P", 1)
    area = ico.volumes
    M00 = mass_matrix(dp0, dp0).toarray()
    np.testing.assert_allclose(M00, np.diag(area), rtol=1e-13, atol=1e-16)
    M11 = mass_matrix(p1, p1)
    assert abs(M11.sum() - area.sum()) < 1e-12 * area.sum()
    # DP1 blocks are single elements: area / 12 [2 1 1; 1 2 1; 1 1 2]
    dp1 = function_space(ico, "DP", 1)
    M = mass_matrix(dp1, dp1).toarray()
    for e in (0, 7, 101):
        d = dp1.local2global[e].astype(int)
        np.testing.assert_allclose(M[np.ix_(d, d)], area[e] / 12.0 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)
    # P1 diagonal: every element at the vertex contributes area / 6
    rows = M11.toarray()
    for v in (0, 5, 40):
        adj = [f for f in range(ico.number_of_elements) if v in p1.local2global[f]]
        assert abs(rows[v, v] - sum(area[f] / 6.0 for f in adj)) < 1e-14
    M01 = mass_matrix(p1, dp0)
    np.testing.assert_allclose(np.asarray(M01.sum(axis=1)).ravel(), area, rtol=1e-13)


def test_identity_operator_weak_form(ico):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary.sparse import identity

    p1 = function_space(ico, "P", 1)
    dp0 = function_space(ico, "DP", 0)
    op = identity(p1, p1, dp0)
    W = op.weak_form()
    assert W.shape == (dp0.global_dof_count, p1.global_dof_count)
    np.testing.assert_allclose(W.to_dense().sum(axis=1), ico.volumes, rtol=1e-13)
    np.testing.assert_allclose((0.5 * W).to_dense(), 0.5 * W.to_dense())


def test_grid_function_callable_projection_and_round_trip(ico):
    import bempp_amd.api as bempp
    from bempp_amd.api import function_space

    p1 = function_space(ico, "P", 1)
    dp0 = function_space(ico, "DP", 0)

    @bempp.real_callable
    def one(x, n, domain_index, res):
        res[0] = 1.0

    @bempp.complex_callable
    def wave(x, n, domain_index, res):
        res[0] = np.exp(1j * 2.0 * x[0]) * (1.0 + n[2])

    f = bempp.GridFunction(dp0, fun=one)
    np.testing.assert_allclose(f.projections(), ico.volumes, rtol=1e-13)
    np.testing.assert_allclose(f.coefficients, np.ones(dp0.global_dof_count), rtol=1e-12)
    np.testing.assert_allclose(f.evaluate_on_element_centers()[0], np.ones(ico.number_of_elements), rtol=1e-12)
    g = bempp.GridFunction(p1, fun=wave)
    c = g.coefficients
    h = bempp.GridFunction(p1, coefficients=c)
    np.testing.assert_allclose(h.projections(p1), g.projections(), rtol=1e-10, atol=1e-14)
    # the vectorized form of the same callable gives the same projections
    vec = bempp.assembly.grid_function.callable(complex=True, vectorized=True)(
        lambda x, n, d, res: res.__setitem__(0, np.exp(1j * 2.0 * x[0]) * (1.0 + n[2])))
    np.testing.assert_allclose(bempp.GridFunction(p1, fun=vec).projections(), g.projections(), rtol=1e-13)
    # integrate and the L2 norm of a constant
    assert abs(bempp.GridFunction(p1, fun=one).integrate()[0] - ico.volumes.sum()) < 1e-10
    assert abs(bempp.GridFunction.from_ones(dp0).l2_norm() - np.sqrt(ico.volumes.sum())) < 1e-12
    with pytest.raises(ValueError):
        bempp.GridFunction(p1)


def test_blocked_operator_space_checks(ico):
    import bempp_amd.api as bempp
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary.sparse import identity

    p1 = function_space(ico, "P", 1)
    dp0 = function_space(ico, "DP", 0)
    blocked = bempp.BlockedOperator(2, 2)
    blocked[0, 0] = identity(p1, p1, p1)
    with pytest.raises(ValueError):
        blocked[0, 1] = identity(dp0, dp0, dp0)  # dual space of row 0 is P1
    assert blocked.domain_spaces[0] == p1 and blocked[1, 1] is None


def test_inverse_mass_matrix_square_thin_thick(ico):
    """Square: M^{-1}; thin / thick: the normal-equations pseudo-inverse (least squares)."""
    from bempp_amd.api import function_space
    from bempp_amd.api.utils.helpers import get_inverse_mass_matrix, get_mass_matrix

    p1 = function_space(ico, "P", 1)
    dp0 = function_space(ico, "DP", 0)
    rng = np.random.default_rng(2)
    for dom, dual in ((p1, p1), (p1, dp0), (dp0, p1)):
        M = get_mass_matrix(dom, dual).toarray()
        b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
        x = get_inverse_mass_matrix(dom, dual) @ b
        np.testing.assert_allclose(x, np.linalg.pinv(M) @ b, rtol=1e-9, atol=1e-12 * np.abs(x).max())
