"""Grid functions (bempp_cl/api/assembly/grid_function.py:12-660).

A GridFunction is given by coefficients in its space, by projections onto a dual space, or by
a Python callable that is projected onto the dual space with the regular triangle rule. The
two representations are converted with the mass matrix of (space, dual_space), as in the
reference: projections = M c, coefficients = M^{-1} projections. All of this is O(F) host work
in numpy/scipy; the callables run in Python (the reference JIT-compiles them with numba, which
this library does not use, so `jit` is accepted and ignored).

Scalar spaces (DP0, DP1, P1) are supported for callables, projections and evaluation;
coefficient-only grid functions of any space (for example RWG densities of potentials) work as
before.
"""

import numpy as _np


def callable(*args, complex=False, jit=True, parameterized=False, vectorized=False):
    """Mark a Python function as a grid-function callable (grid_function.py:12-90).

    Non-vectorized: f(x, n, domain_index, result[, parameters]) for one point x (3,), normal n
    (3,); vectorized: the same with x, n of shape (3, npoints) and result (dim, npoints).
    """

    def wrap(f):
        if parameterized:
            def wrapper_callable(x, n, domain_index, res, parameters):
                f(x, n, domain_index, res, parameters)
        else:
            def wrapper_callable(x, n, domain_index, res, parameters):
                f(x, n, domain_index, res)
        wrapper_callable.bempp_type = "complex" if complex else "real"
        wrapper_callable.bempp_vectorized = bool(vectorized)
        return wrapper_callable

    if not args:
        return wrap
    return wrap(args[0])


def real_callable(*args, jit=True):
    """Wrap a function as a real callable."""
    return callable(*args, complex=False, jit=jit)


def complex_callable(*args, jit=True):
    """Wrap a function as a complex callable."""
    return callable(*args, complex=True, jit=jit)


def _scalar_only(space, what):
    if space.codomain_dimension != 1:
        raise NotImplementedError(f"{what} is provided for scalar spaces (DP0, DP1, P1).")


def _project_function(fun, dual_space, order, function_parameters):
    """Projections sum_e J_e sum_q w_q f(x_q) phi_i(x_q) onto the dual space (grid_function.py:744-880)."""
    from bempp_amd.api.integration.triangle_gauss import rule

    _scalar_only(dual_space, "Projecting a callable")
    points, weights = rule(order)
    grid = dual_space.grid
    dtype = _np.complex128 if getattr(fun, "bempp_type", "real") == "complex" else _np.float64
    params = _np.array([], dtype=_np.float64) if function_parameters is None else _np.asarray(function_parameters)
    if dtype == _np.complex128:
        params = params.astype(_np.complex128)
    elements = dual_space.support_elements
    nq = points.shape[1]
    verts = grid.vertices
    el = grid.elements[:, elements]
    bary = _np.vstack([1.0 - points[0] - points[1], points[0], points[1]])  # (3, nq)
    # quadrature points (3, nelements, nq) and normals with the space's normal multipliers
    xq = _np.einsum("dve,vq->deq", verts[:, el], bary)
    normals = (grid.normals[elements] * dual_space.normal_multipliers[elements][:, None]).T  # (3, nelements)
    domains = grid.domain_indices[elements]
    values = _np.empty((1, len(elements), nq), dtype=dtype)
    if getattr(fun, "bempp_vectorized", False):
        flat = _np.empty((1, len(elements) * nq), dtype=dtype)
        fun(xq.reshape(3, -1), _np.repeat(normals, nq, axis=1), _np.repeat(domains, nq).astype(_np.float64), flat,
            params)
        values[:] = flat.reshape(1, len(elements), nq)
    else:
        res = _np.empty(1, dtype=dtype)
        for k in range(len(elements)):
            n = _np.ascontiguousarray(normals[:, k])
            for q in range(nq):
                fun(_np.ascontiguousarray(xq[:, k, q]), n, domains[k], res, params)
                values[0, k, q] = res[0]
    phi = dual_space.shapeset.evaluate(points)[0]  # (nshape, nq)
    local = _np.einsum("eq,iq,q->ei", values[0], phi, weights)
    local = local * grid.integration_elements[elements][:, None] * dual_space.local_multipliers[elements]
    out = _np.zeros(dual_space.global_dof_count, dtype=dtype)
    _np.add.at(out, dual_space.local2global[elements].astype(_np.int64), local)
    return out


class GridFunction(object):
    """A function on a grid, in a space (grid_function.py:106-660)."""

    def __init__(self, space, dual_space=None, fun=None, coefficients=None, projections=None, parameters=None,
                 function_parameters=None):
        import bempp_amd.api as api

        if dual_space is None:
            dual_space = space
        if space.grid != dual_space.grid or not _np.all(space.normal_multipliers == dual_space.normal_multipliers):
            raise ValueError("Space and dual space must be defined on the same grid with same normal directions.")
        if sum(1 for e in (fun, coefficients, projections) if e is not None) != 1:
            raise ValueError("Exactly one of 'fun', 'coefficients' or 'projections' must be nonzero.")
        self._space = space
        self._dual_space = dual_space
        self._parameters = api.assign_parameters(parameters)
        self._coefficients = None
        self._projections = None
        if coefficients is not None:
            coefficients = _np.asarray(coefficients)
            if coefficients.shape != (space.global_dof_count,):
                raise ValueError("coefficients must have one entry per global dof")
            self._coefficients = coefficients
            self._representation = "primal"
        elif projections is not None:
            projections = _np.asarray(projections)
            if projections.shape != (dual_space.global_dof_count,):
                raise ValueError("projections must have one entry per global dof of the dual space")
            self._projections = projections
            self._representation = "dual"
        else:
            self._projections = _project_function(fun, dual_space, self._parameters.quadrature.regular,
                                                   function_parameters)
            self._representation = "dual"

    @property
    def space(self):
        return self._space

    @property
    def dual_space(self):
        return self._dual_space

    @property
    def parameters(self):
        return self._parameters

    @property
    def representation(self):
        """'primal' if given by coefficients, 'dual' if given by projections."""
        return self._representation

    @property
    def dtype(self):
        if self._coefficients is not None:
            return self._coefficients.dtype
        return self._projections.dtype

    @property
    def component_count(self):
        return self._space.codomain_dimension

    @property
    def coefficients(self):
        """Coefficient vector (from the projections with the inverse mass matrix if needed)."""
        if self._coefficients is None:
            from bempp_amd.api.utils.helpers import get_inverse_mass_matrix

            self._coefficients = get_inverse_mass_matrix(self._space, self._dual_space) @ self._projections
            self._representation = "primal"
        return self._coefficients

    @property
    def grid_coefficients(self):
        """Coefficients per grid dof (this library's spaces need no dof transformation)."""
        return self.coefficients

    def projections(self, dual_space=None):
        """Vector of projections onto dual_space (default: the grid function's dual space)."""
        from bempp_amd.api.utils.helpers import get_mass_matrix

        if dual_space is None:
            dual_space = self._dual_space
        if dual_space == self._dual_space and self._projections is not None:
            return self._projections
        return get_mass_matrix(self._space, dual_space) @ self.coefficients

    def evaluate(self, element_index, local_coordinates):
        """Values (component_count, npoints) on one element at reference points (2, npoints)."""
        _scalar_only(self._space, "Evaluation")
        dofs = self._space.local2global[element_index]
        vals = self._space.evaluate(element_index, local_coordinates)  # (1, nshape, npoints)
        return _np.tensordot(vals, self.grid_coefficients[dofs], axes=([1], [0]))

    def evaluate_on_element_centers(self):
        """Values (component_count, number_of_elements) at the element centroids (zero off the support)."""
        local = _np.array([[1.0 / 3], [1.0 / 3]])
        values = _np.zeros((self.component_count, self._space.grid.number_of_elements), dtype=self.dtype)
        for index in self._space.support_elements:
            values[:, index] = self.evaluate(index, local).flat
        return values

    def evaluate_on_vertices(self):
        """Values at the vertices: area-weighted averages of the adjacent supported elements."""
        grid = self._space.grid
        local = _np.array([[0, 1, 0], [0, 0, 1]], dtype="float64")
        values = _np.zeros((self.component_count, grid.number_of_vertices), dtype=self.dtype)
        areas = _np.zeros(grid.number_of_vertices)
        used = _np.zeros(grid.number_of_vertices, dtype=bool)
        for e in self._space.support_elements:
            v = self.evaluate(e, local)
            for i in range(3):
                idx = grid.elements[i, e]
                used[idx] = True
                areas[idx] += grid.volumes[e]
                values[:, idx] += v[:, i] * grid.volumes[e]
        values[:, used] /= areas[used]
        return values

    def integrate(self):
        """Integral of the function over its support (component_count,)."""
        from bempp_amd.api.integration.triangle_gauss import rule

        _scalar_only(self._space, "Integration")
        points, weights = rule(self._parameters.quadrature.regular)
        sp = self._space
        el = sp.support_elements
        phi = sp.shapeset.evaluate(points)[0]  # (nshape, nq)
        c = self.grid_coefficients[sp.local2global[el]] * sp.local_multipliers[el]  # (ne, nshape)
        return _np.array([_np.sum((c @ phi) @ weights * sp.grid.integration_elements[el])])

    def l2_norm(self):
        """L2 norm sqrt(c^H M c)."""
        vec = self.coefficients
        return _np.sqrt(_np.abs(vec.conjugate() @ (self._space.mass_matrix() @ vec)))

    def __add__(self, other):
        if self._space != other.space:
            raise ValueError("Spaces are not identical.")
        if self._representation == "dual" and other.representation == "dual" and self._dual_space == other.dual_space:
            return GridFunction(self._space, projections=self.projections() + other.projections(),
                                dual_space=self._dual_space)
        return GridFunction(self._space, coefficients=self.coefficients + other.coefficients)

    def __mul__(self, alpha):
        if not _np.isscalar(alpha):
            return NotImplemented
        if self._representation == "dual":
            return GridFunction(self._space, projections=alpha * self._projections, dual_space=self._dual_space,
                                parameters=self._parameters)
        return GridFunction(self._space, coefficients=alpha * self.coefficients, parameters=self._parameters)

    def __rmul__(self, alpha):
        if _np.isscalar(alpha):
            return self * alpha
        return NotImplemented

    def __truediv__(self, alpha):
        return self * (1.0 / alpha)

    def __neg__(self):
        return self * -1.0

    def __sub__(self, other):
        if self._space != other.space:
            raise ValueError("Spaces are not identical.")
        return self + (-other)

    @classmethod
    def from_random(cls, space):
        """Random coefficients, normalised to unit L2 norm."""
        fun = cls(space, coefficients=_np.random.randn(space.global_dof_count))
        return fun / fun.l2_norm()

    @classmethod
    def from_ones(cls, space):
        return cls(space, coefficients=_np.ones(space.global_dof_count))

    @classmethod
    def from_zeros(cls, space):
        return cls(space, coefficients=_np.zeros(space.global_dof_count))


__all__ = ["GridFunction", "callable", "real_callable", "complex_callable"]
