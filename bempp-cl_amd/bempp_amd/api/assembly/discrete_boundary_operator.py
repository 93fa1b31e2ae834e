"""Dense discrete operators kept in HBM (bempp_cl/api/assembly/discrete_boundary_operator.py:177-252).

The reference wraps a host numpy array in a scipy LinearOperator. Here the matrix stays a
device tensor (`.tensor`); `.A` / `to_dense()` copy it to host memory once and cache it,
matrix-vector products run on the device.
"""

import numpy as _np


class DenseDiscreteBoundaryOperator(object):
    """Dense matrix of a boundary operator, resident on the GPU."""

    def __init__(self, tensor):
        self._tensor = tensor
        self._host = None

    @property
    def tensor(self):
        """The (rows, cols) device tensor."""
        return self._tensor

    @property
    def shape(self):
        return tuple(self._tensor.shape)

    @property
    def dtype(self):
        return _np.dtype(str(self._tensor.dtype).replace("torch.", ""))

    @property
    def A(self):
        """Host numpy copy of the matrix."""
        if self._host is None:
            self._host = self._tensor.cpu().numpy()
        return self._host

    def to_dense(self):
        return self.A

    def matvec(self, x):
        import torch

        squeeze = _np.ndim(x) == 1
        xt = torch.as_tensor(_np.asarray(x), device=self._tensor.device)
        if xt.dtype != self._tensor.dtype:
            if xt.is_complex() and not self._tensor.is_complex():
                re = self.matvec(_np.real(x))
                im = self.matvec(_np.imag(x))
                return re + 1j * im
            xt = xt.to(self._tensor.dtype)
        if squeeze:
            xt = xt[:, None]
        y = (self._tensor @ xt).cpu().numpy()
        return y[:, 0] if squeeze else y

    def __matmul__(self, x):
        return self.matvec(x)

    def dot(self, x):
        return self.matvec(x)

    def __mul__(self, alpha):
        if _np.isscalar(alpha):
            return DenseDiscreteBoundaryOperator(self._tensor * alpha)
        return self.matvec(alpha)

    def __add__(self, other):
        return DenseDiscreteBoundaryOperator(self._tensor + other._tensor)

    def transpose(self):
        return DenseDiscreteBoundaryOperator(self._tensor.T.contiguous())
