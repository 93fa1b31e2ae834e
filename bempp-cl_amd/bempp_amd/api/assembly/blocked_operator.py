"""Blocked operators (bempp_cl/api/assembly/blocked_operator.py:25-236, 707-790).

An m x n BlockedOperator holds boundary operators whose rows share range and dual spaces and
whose columns share domain spaces. Its weak form is one dense device matrix with the blocks
in place (empty blocks are zero), so that solvers apply it with a single HIP product.
"""

import numpy as _np


class BlockedOperator(object):
    """An m x n blocked boundary operator."""

    def __init__(self, m, n):
        self._operators = _np.empty((m, n), dtype=object)
        self._rows = m * [False]
        self._cols = n * [False]
        self._dual_to_range_spaces = m * [None]
        self._range_spaces = m * [None]
        self._domain_spaces = n * [None]
        self._cached = None

    @property
    def ndims(self):
        return self._operators.shape

    def __getitem__(self, key):
        return self._operators[key]

    def __setitem__(self, key, operator):
        row, col = key
        if self._range_spaces[row] is not None and operator.range != self._range_spaces[row]:
            raise ValueError(f"Range space not compatible with self.range_spaces[{row}]")
        if self._dual_to_range_spaces[row] is not None and operator.dual_to_range != self._dual_to_range_spaces[row]:
            raise ValueError(f"Dual to range space not compatible with self.dual_to_range_spaces[{row}]")
        if self._domain_spaces[col] is not None and operator.domain != self._domain_spaces[col]:
            raise ValueError(f"Domain space not compatible with self.domain_spaces[{col}]")
        self._range_spaces[row] = operator.range
        self._dual_to_range_spaces[row] = operator.dual_to_range
        self._domain_spaces[col] = operator.domain
        self._operators[key] = operator
        self._rows[row] = True
        self._cols[col] = True
        self._cached = None

    @property
    def range_spaces(self):
        return tuple(self._range_spaces)

    @property
    def dual_to_range_spaces(self):
        return tuple(self._dual_to_range_spaces)

    @property
    def domain_spaces(self):
        return tuple(self._domain_spaces)

    def weak_form(self):
        """The blocked weak form (assembled on first use)."""
        if self._cached is None:
            if False in self._rows or False in self._cols:
                raise ValueError("Each row and column must have at least one operator")
            m, n = self.ndims
            ops = _np.empty((m, n), dtype=object)
            for i in range(m):
                for j in range(n):
                    if self._operators[i, j] is not None:
                        ops[i, j] = self._operators[i, j].weak_form()
            self._cached = BlockedDiscreteOperator(
                ops, [s.global_dof_count for s in self._dual_to_range_spaces],
                [s.global_dof_count for s in self._domain_spaces])
        return self._cached

    def __mul__(self, other):
        """Apply to a list of grid functions (one per column)."""
        from bempp_amd.api.assembly.grid_function import GridFunction

        if not isinstance(other, (list, tuple)):
            return NotImplemented
        x = coefficients_from_grid_functions_list(other)
        y = self.weak_form() @ x
        out, pos = [], 0
        for rng, dual in zip(self._range_spaces, self._dual_to_range_spaces):
            k = dual.global_dof_count
            out.append(GridFunction(rng, projections=y[pos:pos + k], dual_space=dual))
            pos += k
        return out


class BlockedDiscreteOperator(object):
    """Discrete blocked operator: the blocks copied into one dense device matrix."""

    def __init__(self, ops, row_sizes, col_sizes):
        import torch

        from bempp_amd.api.assembly.discrete_boundary_operator import DenseDiscreteBoundaryOperator, _torch_dtype
        from bempp_amd.core import hip_backend as hb

        dtype = _np.dtype("float32")
        for op in ops.flat:
            if op is not None:
                dtype = _np.promote_types(dtype, op.dtype)
        self._row_offsets = _np.concatenate([[0], _np.cumsum(row_sizes)])
        self._col_offsets = _np.concatenate([[0], _np.cumsum(col_sizes)])
        full = torch.zeros((int(self._row_offsets[-1]), int(self._col_offsets[-1])), dtype=_torch_dtype(dtype),
                           device=hb.device())
        for i in range(ops.shape[0]):
            for j in range(ops.shape[1]):
                if ops[i, j] is not None:
                    r0, r1 = self._row_offsets[i], self._row_offsets[i + 1]
                    c0, c1 = self._col_offsets[j], self._col_offsets[j + 1]
                    full[r0:r1, c0:c1] = ops[i, j].device_tensor(dtype)
        self._ops = ops
        self._dense = DenseDiscreteBoundaryOperator(full)

    @property
    def tensor(self):
        return self._dense.tensor

    @property
    def shape(self):
        return self._dense.shape

    @property
    def dtype(self):
        return self._dense.dtype

    def __getitem__(self, key):
        return self._ops[key]

    def device_tensor(self, dtype=None):
        return self._dense.device_tensor(dtype)

    def to_dense(self):
        return self._dense.to_dense()

    @property
    def A(self):
        return self._dense.A

    def matvec(self, x):
        return self._dense.matvec(x)

    def __matmul__(self, x):
        return self.matvec(x)

    def dot(self, x):
        return self.matvec(x)


def coefficients_from_grid_functions_list(grid_funs):
    """One vector with the coefficients of every grid function in the list."""
    dtype = _np.dtype("float32")
    for item in grid_funs:
        dtype = _np.promote_types(dtype, item.coefficients.dtype)
    return _np.concatenate([_np.asarray(item.coefficients, dtype=dtype) for item in grid_funs])


def projections_from_grid_functions_list(grid_funs, projection_spaces):
    """One vector with the projections of grid_funs[i] onto projection_spaces[i]."""
    projections = [item.projections(space) for item, space in zip(grid_funs, projection_spaces)]
    dtype = _np.dtype("float32")
    for p in projections:
        dtype = _np.promote_types(dtype, p.dtype)
    return _np.concatenate([_np.asarray(p, dtype=dtype) for p in projections])


def grid_function_list_from_coefficients(coefficients, spaces):
    """Split a coefficient vector into grid functions of the given spaces."""
    from bempp_amd.api.assembly.grid_function import GridFunction

    out, pos = [], 0
    for space in spaces:
        k = space.global_dof_count
        out.append(GridFunction(space, coefficients=coefficients[pos:pos + k]))
        pos += k
    return out
