from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR
"""Boundary operators with cached weak forms (bempp_cl/api/assembly/boundary_operator.py:4-131)."""


class BoundaryOperator(object):
    """Base class of boundary operators."""

    def __init__(self, domain, range_, dual_to_range, parameters):
        self._domain = domain
        self._range = range_
        self._dual_to_range = dual_to_range
        self._parameters = parameters
        self._cached = None

    @property
    def domain(self):
        return self._domain

    @property
    def range(self):
        return self._range

    @property
    def dual_to_range(self):
        return self._dual_to_range

    @property
    def parameters(self):
        return self._parameters

    def weak_form(self):
        """Return the discrete weak form, assembling it on first use."""
        if self._cached is None:
            self._cached = self._assemble()
        return self._cached

    def __mul__(self, other):
        import numpy as np

        if np.isscalar(other):
            return _ScaledBoundaryOperator(self, other)
        if isinstance(other, BoundaryOperator):
            raise NotImplementedError("Products of boundary operators are not provided.")
        from bempp_amd.api.assembly.grid_function import GridFunction

        if isinstance(other, GridFunction):
            if not self.domain.is_compatible(other.space):
                raise ValueError("Operator domain space does not match GridFunction space.")
            return GridFunction(self.range, projections=self.weak_form() @ other.coefficients,
                                dual_space=self.dual_to_range)
        return NotImplemented

    def __rmul__(self, other):
        import numpy as np

        if np.isscalar(other):
            return _ScaledBoundaryOperator(self, other)
        return NotImplemented

    def __add__(self, other):
        return _SumBoundaryOperator(self, other)

    def __neg__(self):
        return self.__mul__(-1.0)

    def __sub__(self, other):
        return self.__add__(-other)


class BoundaryOperatorWithAssembler(BoundaryOperator):
    """A boundary operator assembled by an AssemblerInterface (boundary_operator.py:99-131)."""

    def __init__(self, domain, range_, dual_to_range, assembler, operator_descriptor):
        super().__init__(domain, range_, dual_to_range, assembler.parameters)
        self._assembler = assembler
        self._operator_descriptor = operator_descriptor

    @property
    def assembler(self):
        return self._assembler

    @property
    def descriptor(self):
        return self._operator_descriptor

    def _assemble(self):
        return self.assembler.assemble(self.descriptor)


class _SumBoundaryOperator(BoundaryOperator):
    def __init__(self, op1, op2):
        if not (op1.domain.is_compatible(op2.domain) and op1.dual_to_range.is_compatible(op2.dual_to_range)):
            raise ValueError("Spaces not compatible.")
        super().__init__(op1.domain, op1.range, op1.dual_to_range, None)
        self._op1, self._op2 = op1, op2

    def _assemble(self):
        return self._op1.weak_form() + self._op2.weak_form()


class _ScaledBoundaryOperator(BoundaryOperator):
    def __init__(self, op, alpha):
        super().__init__(op.domain, op.range, op.dual_to_range, op.parameters)
        self._op, self._alpha = op, alpha

    def _assemble(self):
        return self._op.weak_form() * self._alpha


def _fusion_key(op):
    """Settings that must agree for two operators to be assembled in one pass."""
    p = op.assembler.parameters
    hip = getattr(p, "hip", None)
    return (
        p.quadrature.regular,
        p.quadrature.singular,
        p.assembly.always_promote_to_double,
        getattr(hip, "mixed_near_field_factor", DEFAULT_NEAR_FIELD_FACTOR),
        getattr(hip, "kahan", False),
        getattr(hip, "symmetric", True),
        op.assembler._device_interface,
    )


def assemble_weak_forms(operators):
    """weak_form() of every operator in the list, in order.

    A single layer and the double layer of the same family (Laplace or Helmholtz) with the same
    spaces, wavenumber, precision and quadrature, both with the dense assembler, are assembled
    together in one pass that computes every point pair's distance and e^{ikr} once (the
    matrices are those weak_form() returns). The results are cached on the operators as
    weak_form() caches them.
    """
    from bempp_amd.core.dense_assembler import DenseAssembler, can_fuse_sl_dl, hip_dense_assemble_sl_dl
    from bempp_amd.core.dispatcher import check_device_interface

    ops = list(operators)

    def dense(op):
        return (isinstance(op, BoundaryOperatorWithAssembler) and op._cached is None
                and isinstance(op.assembler._implementation, DenseAssembler))

    for i, a in enumerate(ops):
        if not dense(a):
            continue
        for b in ops[i + 1:]:
            if not dense(b) or _fusion_key(a) != _fusion_key(b):
                continue
            for sl, dl in ((a, b), (b, a)):
                if can_fuse_sl_dl((sl.descriptor, sl.domain, sl.dual_to_range), (dl.descriptor, dl.domain, dl.dual_to_range)):
                    check_device_interface(sl.assembler._device_interface)
                    m1, m2 = hip_dense_assemble_sl_dl(sl.domain, sl.dual_to_range, sl.assembler.parameters,
                                                      sl.descriptor, dl.descriptor)
                    sl._cached = _wrap(m1, sl.assembler.parameters)
                    dl._cached = _wrap(m2, dl.assembler.parameters)
                    break
            if a._cached is not None:
                break
    return [op.weak_form() for op in ops]


def _wrap(mat, parameters):
    import torch
    from bempp_amd.api.assembly.discrete_boundary_operator import DenseDiscreteBoundaryOperator

    if parameters.assembly.always_promote_to_double:
        if mat.dtype == torch.float32:
            mat = mat.double()
        elif mat.dtype == torch.complex64:
            mat = mat.to(torch.complex128)
    return DenseDiscreteBoundaryOperator(mat)
