"""Differentiable single-layer operator application and potential (torch.autograd).

There is no counterpart in the reference, which assembles matrices but has no derivatives.
Two torch.autograd.Function operations on a DP0 space over a flat triangle mesh:

* `single_layer_apply(vertices, coefficients, problem)` -> V(vertices) c, the dense Galerkin
  single layer (Helmholtz e^{ikr}/(4 pi r), or Laplace for k = 0) applied to c. The forward
  assembles V in HBM (mixed FP32/FP64 or FP64, as `weak_form()`) and applies it; the backward
  returns V^H g for c and the vertex gradient from the adjoint kernel, which recomputes the
  element geometry and the kernel derivatives pair by pair instead of storing them.
* `single_layer_potential(vertices, coefficients, points, problem)` -> P(vertices) c, the
  single-layer potential at the given points; its backward returns P^H g and the vertex
  gradient (FP64 adjoint kernel).

`vertices` is a (n_vertices, 3) float64 tensor, `coefficients` a complex tensor of one
coefficient per element. Gradients follow torch's convention for complex tensors (for a real
loss L of a complex output y, the backward receives dL/d(conj y) times 2).
"""

from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR
import numpy as _np


class SingleLayerProblem(object):
    """Mesh topology and operator settings shared by the differentiable operations.

    elements: (3, n_elements) vertex indices; wavenumber: real k (0 = Laplace);
    precision: "single" (mixed FP32/FP64) or "double".
    """

    def __init__(self, elements, wavenumber, precision="single", parameters=None):
        import bempp_amd.api as api
        from bempp_amd.api.grid.grid import Grid

        self.elements = _np.asarray(elements)
        if _np.iscomplexobj(wavenumber) and _np.imag(wavenumber) != 0:
            raise ValueError("the differentiable single layer takes a real wavenumber")
        self.wavenumber = float(_np.real(wavenumber))
        if precision not in ("single", "double"):
            raise ValueError("precision must be 'single' or 'double'")
        self.precision = precision
        self.parameters = api.assign_parameters(parameters)
        if self.parameters.quadrature.regular != 4:
            raise ValueError("the adjoint kernels use the order-4 (6-point) regular rule")
        self._topology = None
        self._Grid = Grid
        self._last = None
        self._space = None

    def grid(self, vertices):
        """Grid of this topology at the given (n_vertices, 3) positions (tensor or array)."""
        import torch

        v = vertices.detach().cpu().numpy() if torch.is_tensor(vertices) else _np.asarray(vertices)
        v = _np.asarray(v, dtype=_np.float64)
        if v.ndim != 2 or v.shape[1] != 3:
            raise ValueError("vertices must have shape (n_vertices, 3)")
        if self._last is not None and self._last[0].shape == v.shape and _np.array_equal(self._last[0], v):
            return self._last[1]
        if self._topology is None:
            self._topology = self._Grid(v.T, self.elements)
            g = self._topology
        else:
            g = self._topology.with_vertices(v.T)
        if self._last is not None:
            # the previous vertex set is superseded: free its device copies (geometry, dof maps,
            # near-field candidates) so that a loop over moving vertices does not accumulate them
            from bempp_amd.core import release

            old_space = self._space if self._space is not None and self._space.grid is self._last[1] else None
            release(self._last[1], (old_space,))
            if old_space is not None:
                self._space = None
        self._last = (v.copy(), g)
        return g

    def space(self, grid):
        """The DP0 space of a grid from `grid` (kept for the last grid, so device copies are reused)."""
        if self._space is None or self._space.grid is not grid:
            if self._space is not None:
                from bempp_amd.core import release

                release(None, (self._space,))
            self._space = _dp0_space(grid)
        return self._space

    def operator(self, space):
        from bempp_amd.api.operators.boundary import helmholtz, laplace

        if self.wavenumber == 0.0:
            return laplace.single_layer(space, space, space, parameters=self.parameters, precision=self.precision)
        return helmholtz.single_layer(space, space, space, self.wavenumber, parameters=self.parameters,
                                      precision=self.precision)

    def potential(self, space, points):
        from bempp_amd.api.operators.potential import helmholtz, laplace

        if self.wavenumber == 0.0:
            return laplace.single_layer(space, points, parameters=self.parameters, precision=self.precision)
        return helmholtz.single_layer(space, points, self.wavenumber, parameters=self.parameters,
                                      precision=self.precision)

    def mode(self):
        from bempp_amd.core import hip_backend as hb

        return hb.PRECISION_MIXED if self.precision == "single" else hb.PRECISION_FP64

    def near_factor(self):
        hip = getattr(self.parameters, "hip", None)
        return getattr(hip, "mixed_near_field_factor", DEFAULT_NEAR_FIELD_FACTOR) if hip is not None else DEFAULT_NEAR_FIELD_FACTOR


def _full_coefficients(space, c):
    """Per-element coefficients m_e c_{dof(e)} of a DP0 space (device, complex128)."""
    import torch

    dev = c.device
    l2g = torch.from_numpy(_np.ascontiguousarray(space.local2global[:, 0].astype(_np.int64))).to(dev)
    mult = torch.from_numpy(_np.ascontiguousarray(space.local_multipliers[:, 0])).to(dev)
    return (c[l2g] * mult).contiguous()


def _device_rule(parameters):
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb

    qp, qw = triangle_rule(parameters.quadrature.regular)
    return hb.quad_rule(qp, qw)


def _dp0_space(grid):
    from bempp_amd.api.space.space import function_space

    return function_space(grid, "DP", 0)


def _make_functions():
    import torch

    class _SingleLayerApply(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vertices, coefficients, problem):
            from bempp_amd.core import hip_backend as hb
            from bempp_amd.core.dense_assembler import hip_dense_assemble

            grid = problem.grid(vertices)
            space = problem.space(grid)
            dev = hb.device()
            c = coefficients.detach().to(dev, torch.complex128).contiguous()
            if c.shape != (space.global_dof_count,):
                raise ValueError(f"coefficients must have shape ({space.global_dof_count},)")
            A = hip_dense_assemble(space, space, problem.parameters, problem.operator(space).descriptor)
            y = hb.dense_matvec(A, c)
            ctx.problem, ctx.space, ctx.A, ctx.c = problem, space, A, c
            ctx.vshape, ctx.vdevice, ctx.cdevice = tuple(vertices.shape), vertices.device, coefficients.device
            return y

        @staticmethod
        def backward(ctx, g):
            from bempp_amd.core import hip_backend as hb
            from bempp_amd.core.near_field import device_near_candidates
            from bempp_amd.core.singular_assembler import singular_rule

            problem, space = ctx.problem, ctx.space
            g = g.detach().to(ctx.A.device, torch.complex128).contiguous()
            grad_v = grad_c = None
            if ctx.needs_input_grad[1]:
                grad_c = hb.dense_matvec(ctx.A, g, trans=2).to(ctx.cdevice)
            if ctx.needs_input_grad[0]:
                grad = torch.zeros((space.grid.number_of_vertices, 3), dtype=torch.float64, device=ctx.A.device)
                near = problem.near_factor()
                mode = problem.mode()
                cand = device_near_candidates(space, space, near) if mode == hb.PRECISION_MIXED else None
                srule = singular_rule(space.grid, problem.parameters.quadrature.singular, space.support,
                                      space.support)
                hb.adjoint_single_layer(hb.device_space(space), _device_rule(problem.parameters), problem.wavenumber,
                                        mode, near, g, ctx.c, grad, near_pairs=cand, singular=srule)
                grad_v = grad.to(ctx.vdevice)
            ctx.A = None
            return grad_v, grad_c, None

    class _SingleLayerPotential(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vertices, coefficients, points, problem):
            from bempp_amd.core import hip_backend as hb
            from bempp_amd.core.dense_potential_assembler import hip_potential_assembler

            grid = problem.grid(vertices)
            space = problem.space(grid)
            dev = hb.device()
            c = coefficients.detach().to(dev, torch.complex128).contiguous()
            if c.shape != (space.global_dof_count,):
                raise ValueError(f"coefficients must have shape ({space.global_dof_count},)")
            pts = points.detach().to(dev, torch.float64).contiguous()
            if pts.dim() != 2 or pts.shape[1] != 3:
                raise ValueError("points must have shape (n_points, 3)")
            points_host = pts.cpu().numpy().T
            pot = problem.potential(space, points_host)
            evaluate = hip_potential_assembler(space, pot.descriptor, points_host, problem.parameters)
            out, _ = evaluate(_full_coefficients(space, c), as_tensor=True)
            ctx.problem, ctx.space, ctx.c, ctx.pts = problem, space, c, pts
            ctx.vdevice, ctx.cdevice = vertices.device, coefficients.device
            return out[0].to(torch.complex128)

        @staticmethod
        def backward(ctx, g):
            from bempp_amd.core import hip_backend as hb

            problem, space = ctx.problem, ctx.space
            g = g.detach().to(ctx.c.device, torch.complex128).contiguous()
            grad_c = torch.zeros_like(ctx.c) if ctx.needs_input_grad[1] else None
            grad = None
            if ctx.needs_input_grad[0]:
                grad = torch.zeros((space.grid.number_of_vertices, 3), dtype=torch.float64, device=ctx.c.device)
            if grad_c is not None or grad is not None:
                hb.adjoint_potential_single_layer(hb.device_space(space), _device_rule(problem.parameters),
                                                  problem.wavenumber, ctx.c, ctx.pts, g, grad_coefficients=grad_c,
                                                  grad_vertices=grad)
            return (None if grad is None else grad.to(ctx.vdevice),
                    None if grad_c is None else grad_c.to(ctx.cdevice), None, None)

    return _SingleLayerApply, _SingleLayerPotential


_FUNCS = None


def _funcs():
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _make_functions()
    return _FUNCS


def single_layer_apply(vertices, coefficients, problem):
    """V(vertices) c as a complex128 device tensor, differentiable in vertices and c."""
    return _funcs()[0].apply(vertices, coefficients, problem)


def single_layer_potential(vertices, coefficients, points, problem):
    """Single-layer potential of c at points ((n_points, 3)), differentiable in vertices and c."""
    return _funcs()[1].apply(vertices, coefficients, points, problem)
