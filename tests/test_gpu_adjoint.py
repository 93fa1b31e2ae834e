"""Adjoint (gradient) kernels and the differentiable single layer on the GPU.

Checks, on perturbed icospheres (smooth random radial perturbation, so that no symmetry of
the sphere hides a wrong sign or index):
* the dense matvec (A x, A^T x, A^H x) against torch products;
* the FP64 vertex gradient of Re(x^H V y) against central finite differences (h = 1e-6) of
  the FP64 GPU assembly and of the FP64 CPU oracle;
* the mixed-precision gradient against the FP64 gradient;
* the potential adjoint: P^H g through the identity Re(g^H P d) = Re(grad_c^H d), and its
  vertex gradient against finite differences of the FP64 potential;
* torch.autograd through V c and the potential, against finite differences of the loss
  ||V c - b||^2 + ||u||^2.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _complex(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


# (rows, cols, leading dimension): one partial column chunk and a column tail; several chunks of
# the tiled product; a leading dimension whose rows are not 16-byte aligned (row-per-wave product)
MATVEC_SHAPES = [(300, 257, 280), (100, 4500, 4500), (70, 2301, 2303)]


@pytest.mark.parametrize("shape", MATVEC_SHAPES, ids=[f"{r}x{c}_ld{ld}" for r, c, ld in MATVEC_SHAPES])
@pytest.mark.parametrize("dtype", ["float32", "float64", "complex64", "complex128"])
def test_dense_matvec(dtype, shape, hip_available):
    import torch
    from bempp_amd.core import hip_backend as hb

    rows, cols, ld = shape
    g = torch.Generator().manual_seed(3)
    dt = getattr(torch, dtype)
    full = torch.randn(rows, ld, generator=g, dtype=torch.float64)
    if dt.is_complex:
        full = full + 1j * torch.randn(rows, ld, generator=g, dtype=torch.float64)
    A = full.to(dt).cuda()[:, :cols]  # leading dimension ld >= cols
    A64 = A.to(torch.complex128)
    x = (torch.randn(cols, dtype=torch.float64, generator=g) + 1j * torch.randn(cols, dtype=torch.float64, generator=g)).cuda()
    xt = (torch.randn(rows, dtype=torch.float64, generator=g) + 1j * torch.randn(rows, dtype=torch.float64, generator=g)).cuda()
    ref = [A64 @ x, A64.T @ xt, A64.conj().T @ xt]
    got = [hb.dense_matvec(A, x, 0), hb.dense_matvec(A, xt, 1), hb.dense_matvec(A, xt, 2)]
    for r, y in zip(ref, got):
        assert (torch.linalg.norm(y - r) / torch.linalg.norm(r)).item() < 1e-13


@pytest.fixture(scope="module")
def pert1():
    from bempp_amd.api.grid.shapes import perturbed_icosphere

    return perturbed_icosphere(1)


def _vtx(grid):
    return np.ascontiguousarray(grid.vertices.T)


def _bilinear_gpu(problem, vertices, x, y):
    """Re(x^H V y) with the FP64 GPU assembly."""
    import torch
    from bempp_amd.api.differentiable import _dp0_space
    from bempp_amd.core.dense_assembler import hip_dense_assemble

    grid = problem.grid(vertices)
    sp = _dp0_space(grid)
    A = hip_dense_assemble(sp, sp, problem.parameters, problem.operator(sp).descriptor).cpu().numpy()
    return float(np.real(np.conj(x) @ (A @ y)))


def _bilinear_oracle(problem, vertices, x, y):
    from bempp_amd.api.differentiable import _dp0_space
    from oracle import oracle as O

    grid = problem.grid(vertices)
    sp = _dp0_space(grid)
    op = "helmholtz_single_layer" if problem.wavenumber != 0 else "laplace_single_layer"
    A = O.assemble(op, grid.vertices, grid.elements, sp, sp, problem.wavenumber)
    return float(np.real(np.conj(x) @ (A @ y)))


def _adjoint_grad(problem, vertices, x, y):
    import torch
    from bempp_amd.api.differentiable import _device_rule, _dp0_space
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.near_field import device_near_candidates
    from bempp_amd.core.singular_assembler import singular_rule

    grid = problem.grid(vertices)
    sp = _dp0_space(grid)
    dev = hb.device()
    grad = torch.zeros((grid.number_of_vertices, 3), dtype=torch.float64, device=dev)
    near = problem.near_factor()
    cand = device_near_candidates(sp, sp, near) if problem.mode() == hb.PRECISION_MIXED else None
    srule = singular_rule(grid, problem.parameters.quadrature.singular, sp.support, sp.support)
    hb.adjoint_single_layer(hb.device_space(sp), _device_rule(problem.parameters), problem.wavenumber, problem.mode(),
                            near, torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev), grad, near_pairs=cand,
                            singular=srule)
    return grad.cpu().numpy()


def _fd(fun, vertices, picks):
    out = []
    for v, d in picks:
        vp = vertices.copy()
        vm = vertices.copy()
        vp[v, d] += H
        vm[v, d] -= H
        out.append((fun(vp) - fun(vm)) / (2 * H))
    return np.array(out)


@pytest.mark.parametrize("k", [0.0, 3.0])
def test_fp64_vertex_gradient_vs_fd(pert1, k, default_parameters, hip_available):
    from bempp_amd.api.differentiable import SingleLayerProblem

    prob = SingleLayerProblem(pert1.elements, k, precision="double", parameters=default_parameters)
    rng = np.random.default_rng(11)
    n = pert1.number_of_elements
    x, y = _complex(rng, n), _complex(rng, n)
    v0 = _vtx(pert1)
    grad = _adjoint_grad(prob, v0, x, y)
    picks = [(v, d) for v in range(pert1.number_of_vertices) for d in range(3)]
    fd = _fd(lambda v: _bilinear_gpu(prob, v, x, y), v0, picks)
    ana = np.array([grad[v, d] for v, d in picks])
    assert rel(ana, fd) < 1e-6


def test_fp64_vertex_gradient_vs_oracle_fd(pert1, default_parameters, hip_available):
    from bempp_amd.api.differentiable import SingleLayerProblem

    prob = SingleLayerProblem(pert1.elements, 2.5, precision="double", parameters=default_parameters)
    rng = np.random.default_rng(12)
    n = pert1.number_of_elements
    x, y = _complex(rng, n), _complex(rng, n)
    v0 = _vtx(pert1)
    grad = _adjoint_grad(prob, v0, x, y)
    picks = [(v, d) for v in rng.choice(pert1.number_of_vertices, 8, replace=False) for d in range(3)]
    fd = _fd(lambda v: _bilinear_oracle(prob, v, x, y), v0, picks)
    ana = np.array([grad[v, d] for v, d in picks])
    assert rel(ana, fd) < 1e-6


@pytest.mark.parametrize("level,k", [(2, 4.0), (3, 8.0)])
def test_mixed_vertex_gradient_vs_fp64(level, k, default_parameters, hip_available):
    from bempp_amd.api.differentiable import SingleLayerProblem
    from bempp_amd.api.grid.shapes import perturbed_icosphere

    grid = perturbed_icosphere(level)
    rng = np.random.default_rng(13)
    n = grid.number_of_elements
    x, y = _complex(rng, n), _complex(rng, n)
    v0 = _vtx(grid)
    g64 = _adjoint_grad(SingleLayerProblem(grid.elements, k, "double", default_parameters), v0, x, y)
    g32 = _adjoint_grad(SingleLayerProblem(grid.elements, k, "single", default_parameters), v0, x, y)
    assert rel(g32, g64) < 1e-5


def test_potential_adjoint(pert1, default_parameters, hip_available):
    import torch
    from bempp_amd.api.differentiable import SingleLayerProblem, _device_rule, _dp0_space
    from bempp_amd.core import hip_backend as hb

    k = 3.0
    prob = SingleLayerProblem(pert1.elements, k, precision="double", parameters=default_parameters)
    rng = np.random.default_rng(14)
    n = pert1.number_of_elements
    pts = rng.standard_normal((50, 3))
    pts = 2.0 * pts / np.linalg.norm(pts, axis=1)[:, None]
    c, g, d = _complex(rng, n), _complex(rng, 50), _complex(rng, n)
    v0 = _vtx(pert1)

    def potential(vertices, coeffs):
        from bempp_amd.api.assembly.grid_function import GridFunction

        sp = _dp0_space(prob.grid(vertices))
        return prob.potential(sp, pts.T).evaluate(GridFunction(sp, coefficients=coeffs))[0]

    dev = hb.device()
    sp = _dp0_space(prob.grid(v0))
    gc = torch.zeros(n, dtype=torch.complex128, device=dev)
    gv = torch.zeros((pert1.number_of_vertices, 3), dtype=torch.float64, device=dev)
    hb.adjoint_potential_single_layer(hb.device_space(sp), _device_rule(default_parameters), k,
                                      torch.from_numpy(c).to(dev), torch.from_numpy(pts).to(dev),
                                      torch.from_numpy(g).to(dev), grad_coefficients=gc, grad_vertices=gv)
    gc = gc.cpu().numpy()
    lhs = np.real(np.conj(g) @ potential(v0, d))
    assert abs(lhs - np.real(np.conj(gc) @ d)) < 1e-12 * max(1.0, abs(lhs))
    picks = [(v, dd) for v in range(0, pert1.number_of_vertices, 3) for dd in range(3)]
    fd = _fd(lambda v: float(np.real(np.conj(g) @ potential(v, c))), v0, picks)
    ana = np.array([gv.cpu().numpy()[v, dd] for v, dd in picks])
    assert rel(ana, fd) < 1e-6


@pytest.mark.parametrize("precision", ["double", "single"])
def test_autograd_loss(precision, default_parameters, hip_available):
    import torch
    from bempp_amd.api.differentiable import SingleLayerProblem, single_layer_apply, single_layer_potential
    from bempp_amd.api.grid.shapes import perturbed_icosphere

    grid = perturbed_icosphere(2)
    k = 4.0
    rng = np.random.default_rng(15)
    n = grid.number_of_elements
    c0, b = _complex(rng, n), _complex(rng, n)
    pts = rng.standard_normal((64, 3))
    pts = torch.from_numpy(2.0 * pts / np.linalg.norm(pts, axis=1)[:, None])
    bt = torch.from_numpy(b).cuda()
    prob = SingleLayerProblem(grid.elements, k, precision=precision, parameters=default_parameters)
    prob64 = SingleLayerProblem(grid.elements, k, precision="double", parameters=default_parameters)

    def loss(p, vertices, coeffs):
        y = single_layer_apply(vertices, coeffs, p)
        u = single_layer_potential(vertices, coeffs, pts, p)
        return (y - bt).abs().pow(2).sum() + u.abs().pow(2).sum()

    v = torch.from_numpy(_vtx(grid)).requires_grad_(True)
    c = torch.from_numpy(c0).requires_grad_(True)
    L = loss(prob, v, c)
    L.backward()
    gv, gc = v.grad.numpy(), c.grad.numpy()

    with torch.no_grad():
        def Lv(vv):
            return float(loss(prob64, torch.from_numpy(vv), torch.from_numpy(c0)))

        picks = [(vi, d) for vi in rng.choice(grid.number_of_vertices, 4, replace=False) for d in range(3)]
        fd = _fd(Lv, _vtx(grid), picks)
        ana = np.array([gv[vi, d] for vi, d in picks])
        tol = 1e-6 if precision == "double" else 1e-4
        assert rel(ana, fd) < tol
        # c-gradient: dL/dt at c + t d equals Re(conj(grad_c) . d) (torch's complex convention)
        d = _complex(rng, n)
        hc = 1e-6
        lp = float(loss(prob64, torch.from_numpy(_vtx(grid)), torch.from_numpy(c0 + hc * d)))
        lm = float(loss(prob64, torch.from_numpy(_vtx(grid)), torch.from_numpy(c0 - hc * d)))
        fdc = (lp - lm) / (2 * hc)
        assert abs(np.real(np.conj(gc) @ d) - fdc) < tol * abs(fdc)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_moving_vertices_do_not_accumulate_device_memory(hip_available, precision):
    """A shape-optimisation loop (new vertex positions every step) keeps device memory flat:
    the problem releases the previous grid's and space's cached device copies."""
    import torch

    import bempp_amd.api as bempp
    from bempp_amd.api.differentiable import SingleLayerProblem, single_layer_apply, single_layer_potential

    grid = bempp.shapes.icosphere(2)
    prob = SingleLayerProblem(grid.elements, 2.0, precision=precision)
    rng = np.random.default_rng(9)
    F = grid.number_of_elements
    c = torch.from_numpy(_complex(rng, F)).cuda()
    pts = torch.from_numpy(3.0 * rng.standard_normal((64, 3))).cuda()
    v0 = np.ascontiguousarray(grid.vertices.T)
    usage = []
    for step in range(5):
        v = torch.from_numpy(v0 * (1.0 + 0.01 * step)).cuda().requires_grad_(True)
        loss = single_layer_apply(v, c, prob).abs().pow(2).sum() + single_layer_potential(v, c, pts, prob).abs().pow(2).sum()
        loss.backward()
        del v, loss
        torch.cuda.synchronize()
        usage.append(torch.cuda.memory_allocated())
    # the first step builds the caches; later steps replace them instead of adding to them
    assert max(usage[1:]) <= usage[1], usage
