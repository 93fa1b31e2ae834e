"""Discrete boundary operators (bempp_cl/api/assembly/discrete_boundary_operator.py:12-252).

The reference wraps a host numpy array (dense) or a scipy sparse matrix in a scipy
LinearOperator. Here a dense weak form stays a device tensor (`.tensor`); `.A` / `to_dense()`
copy it to host memory once and cache it, and matrix-vector products run on the device through
the library's HIP product (bem_dense_matvec, FP64 accumulation). Sparse weak forms (the L2
identity) are host CSR matrices; adding one to a dense operator expands it on the device.
"""

import numpy as _np


def _torch_dtype(np_dtype):
    import torch

    return {
        _np.dtype("float32"): torch.float32,
        _np.dtype("float64"): torch.float64,
        _np.dtype("complex64"): torch.complex64,
        _np.dtype("complex128"): torch.complex128,
    }[_np.dtype(np_dtype)]


def _np_dtype(torch_dtype):
    return _np.dtype(str(torch_dtype).replace("torch.", ""))


def device_matvec(tensor, x):
    """tensor @ x for a device matrix and a device vector, as complex128 (HIP product)."""
    import torch
    from bempp_amd.core import hip_backend as hb

    if x.dtype != torch.complex128:
        x = x.to(torch.complex128)
    return hb.dense_matvec(tensor, x.contiguous())


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
        return _np_dtype(self._tensor.dtype)

    @property
    def A(self):
        """Host numpy copy of the matrix."""
        if self._host is None:
            self._host = self._tensor.cpu().numpy()
        return self._host

    def to_dense(self):
        return self.A

    def device_tensor(self, dtype=None):
        """The matrix on the device, converted to the numpy dtype given (no copy if it matches)."""
        if dtype is None or _torch_dtype(dtype) == self._tensor.dtype:
            return self._tensor
        return self._tensor.to(_torch_dtype(dtype))

    def matvec(self, x):
        """A x for a host vector (n,) or block of vectors (n, m); returns numpy like the reference."""
        import torch

        x = _np.asarray(x)
        squeeze = x.ndim == 1
        xs = x[:, None] if squeeze else x
        out_dtype = _np.result_type(self.dtype, xs.dtype)
        cols = []
        for j in range(xs.shape[1]):
            xt = torch.as_tensor(_np.ascontiguousarray(xs[:, j]).astype(_np.complex128), device=self._tensor.device)
            cols.append(device_matvec(self._tensor, xt).cpu().numpy())
        y = _np.stack(cols, axis=1)
        if not _np.issubdtype(out_dtype, _np.complexfloating):
            y = y.real
        y = y.astype(out_dtype, copy=False)
        return y[:, 0] if squeeze else y

    def __matmul__(self, x):
        return self.matvec(x)

    def dot(self, x):
        return self.matvec(x)

    def __mul__(self, alpha):
        if _np.isscalar(alpha):
            return DenseDiscreteBoundaryOperator(self._tensor * alpha)
        return self.matvec(alpha)

    def __rmul__(self, alpha):
        if _np.isscalar(alpha):
            return self.__mul__(alpha)
        return NotImplemented

    def __add__(self, other):
        dtype = _np.result_type(self.dtype, other.dtype)
        return DenseDiscreteBoundaryOperator(self.device_tensor(dtype) + other.device_tensor(dtype))

    def __radd__(self, other):
        return self.__add__(other)

    def __neg__(self):
        return self.__mul__(-1.0)

    def __sub__(self, other):
        return self.__add__(-other)

    def transpose(self):
        return DenseDiscreteBoundaryOperator(self._tensor.T.contiguous())


class SparseDiscreteBoundaryOperator(object):
    """Sparse matrix (scipy CSR, host) of a local operator (discrete_boundary_operator.py:177-252)."""

    def __init__(self, mat):
        self._mat = mat.tocsr()

    @property
    def A(self):
        """The scipy sparse matrix."""
        return self._mat

    @property
    def shape(self):
        return self._mat.shape

    @property
    def dtype(self):
        return self._mat.dtype

    def to_dense(self):
        return self._mat.toarray()

    def device_tensor(self, dtype=None):
        """The matrix as a dense device tensor of the given numpy dtype."""
        import torch
        from bempp_amd.core import hip_backend as hb

        dtype = self.dtype if dtype is None else _np.dtype(dtype)
        coo = self._mat.tocoo()
        out = torch.zeros(self.shape, dtype=_torch_dtype(dtype), device=hb.device())
        if coo.nnz:
            rows = torch.as_tensor(coo.row.astype(_np.int64), device=out.device)
            cols = torch.as_tensor(coo.col.astype(_np.int64), device=out.device)
            vals = torch.as_tensor(coo.data.astype(dtype), device=out.device)
            out.index_put_((rows, cols), vals, accumulate=True)
        return out

    def matvec(self, x):
        return self._mat @ _np.asarray(x)

    def __matmul__(self, x):
        return self.matvec(x)

    def dot(self, x):
        return self.matvec(x)

    def __mul__(self, alpha):
        if _np.isscalar(alpha):
            return SparseDiscreteBoundaryOperator(self._mat * alpha)
        return self.matvec(alpha)

    def __rmul__(self, alpha):
        if _np.isscalar(alpha):
            return self.__mul__(alpha)
        return NotImplemented

    def __add__(self, other):
        if isinstance(other, SparseDiscreteBoundaryOperator):
            return SparseDiscreteBoundaryOperator(self._mat + other.A)
        return other.__add__(self)

    def __neg__(self):
        return self.__mul__(-1.0)

    def __sub__(self, other):
        return self.__add__(-other)

    def transpose(self):
        return SparseDiscreteBoundaryOperator(self._mat.T)


def as_matrix(operator):
    """Dense host (numpy) matrix of a discrete operator (discrete_boundary_operator.py:255-270)."""
    return operator.to_dense()
