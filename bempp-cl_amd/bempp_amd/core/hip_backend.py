"""ctypes binding of libbempp_hip.so and device-resident copies of grids and spaces.

This is the "hip" device interface: the counterpart of the reference's backend modules
bempp_cl/core/numba_assemblers.py and bempp_cl/core/opencl_assemblers.py. Device memory
and streams come from PyTorch-ROCm; all compute happens in the HIP library. There is no
CPU fallback: if the library is missing or no GPU is visible, the calls raise.
"""

from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR
import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "libbempp_hip.so")
_LIB = None

OPERATORS = {
    "laplace_single_layer": 0,
    "laplace_double_layer": 1,
    "laplace_adjoint_double_layer": 2,
    "helmholtz_single_layer": 3,
    "helmholtz_double_layer": 4,
    "helmholtz_adjoint_double_layer": 5,
    "maxwell_electric_field": 6,
    "laplace_hypersingular": 7,
    "helmholtz_hypersingular": 8,
    "helmholtz_far_field_single_layer": 9,
    "helmholtz_far_field_double_layer": 10,
    "maxwell_electric_far_field": 11,
    "maxwell_magnetic_far_field": 12,
    "maxwell_magnetic_field": 13,
}
SHAPESETS = {"p0_discontinuous": 0, "p1_discontinuous": 1, "rwg0": 2, "snc0": 2}
DTYPES = {"float32": 0, "float64": 1, "complex64": 2, "complex128": 3}
PRECISION_FP64 = 0
PRECISION_MIXED = 1

GEOM_STRIDE = 16

_ERRORS = {1: ValueError, 2: KeyError, 3: RuntimeError, 4: ValueError}


class BemSpace(ctypes.Structure):
    _fields_ = [
        ("n_elements", ctypes.c_int32),
        ("shapeset", ctypes.c_int32),
        ("elements", ctypes.c_void_p),
        ("geometry", ctypes.c_void_p),
        ("vertex_ids", ctypes.c_void_p),
        ("local2global", ctypes.c_void_p),
        ("multipliers", ctypes.c_void_p),
        ("normal_multipliers", ctypes.c_void_p),
    ]


class BemDofBlocks(ctypes.Structure):
    _fields_ = [
        ("n_blocks", ctypes.c_int32),
        ("max_dofs", ctypes.c_int32),
        ("start", ctypes.c_void_p),
        ("dofs", ctypes.c_void_p),
        ("ndofs", ctypes.c_void_p),
        ("slots", ctypes.c_void_p),
        ("incidence", ctypes.c_void_p),
        ("max_incidence", ctypes.c_int32),
    ]


QUAD_MAX_POINTS = 79  # BEM_QUAD_MAX_POINTS: the longest symmetric rule (order 20)
QUAD_FAST_POINTS = 16  # BEM_QUAD_FAST_POINTS: longer rules take the FP64 one-lane-per-pair kernel


class BemQuadRule(ctypes.Structure):
    _fields_ = [
        ("n_points", ctypes.c_int32),
        ("x", ctypes.c_double * QUAD_MAX_POINTS),
        ("y", ctypes.c_double * QUAD_MAX_POINTS),
        ("w", ctypes.c_double * QUAD_MAX_POINTS),
    ]


class BemKBatchConfig(ctypes.Structure):
    _fields_ = [("test_tile", ctypes.c_int32), ("workgroups_per_cu", ctypes.c_int32)]


KBATCH_MAX_WAVENUMBERS = 64
KBATCH_REGULAR = 1
KBATCH_NEAR = 2
KBATCH_SINGULAR = 4
KBATCH_ALL = 7


def library_path():
    return _LIB_PATH


def lib():
    """Load libbempp_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"libbempp_hip.so not found at {_LIB_PATH}; build it with `python bempp-cl_amd/build_hip.py`"
            )
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i = ctypes.c_int
        d = ctypes.c_double
        i64 = ctypes.c_int64
        sz = ctypes.c_size_t
        SP = ctypes.POINTER(BemSpace)
        QP = ctypes.POINTER(BemQuadRule)
        BP = ctypes.POINTER(BemDofBlocks)
        L.bem_regular_assemble.argtypes = [i, SP, SP, i, QP, d, d, i, d, i, P, i64, i, i, P]
        L.bem_regular_assemble.restype = i
        L.bem_regular_assemble_blocked.argtypes = [i, SP, SP, BP, BP, i, QP, d, d, i, d, i, P, i64, i, P]
        L.bem_regular_assemble_blocked.restype = i
        L.bem_near_assemble.argtypes = [i, SP, SP, i, P, i, QP, d, d, d, P, i64, i, P]
        L.bem_near_assemble.restype = i
        L.bem_singular_assemble.argtypes = [i, SP, SP, i, P, P, P, P, d, d, P, i64, i, P]
        L.bem_singular_assemble.restype = i
        L.bem_potential_workspace_size.argtypes = [i, SP, QP, i, i]
        L.bem_potential_workspace_size.restype = sz
        L.bem_potential_evaluate.argtypes = [i, SP, P, i, P, QP, d, d, i, P, i, P, sz, P]
        L.bem_potential_evaluate.restype = i
        L.bem_potential_evaluate_noalloc.argtypes = [i, SP, P, i, P, QP, d, d, i, P, i, P]
        L.bem_potential_evaluate_noalloc.restype = i
        L.bem_fill_zero.argtypes = [P, sz, P]
        L.bem_fill_zero.restype = i
        L.bem_last_error.restype = ctypes.c_char_p
        L.bem_version.restype = ctypes.c_char_p
        L.bem_kbatch_assemble.argtypes = [
            SP, SP, i, QP, i, ctypes.POINTER(ctypes.c_double), i, i, d, ctypes.POINTER(BemKBatchConfig), P, i64, i64,
            i, i, P, i, P, P, P, P, P,
        ]
        L.bem_kbatch_assemble.restype = i
        L.bem_adjoint_single_layer.argtypes = [SP, QP, d, i, d, P, P, i, P, i, P, P, P, P, P, P]
        L.bem_adjoint_single_layer.restype = i
        L.bem_adjoint_potential_single_layer.argtypes = [SP, QP, d, P, i, P, P, P, P, P]
        L.bem_adjoint_potential_single_layer.restype = i
        L.bem_dense_matvec.argtypes = [i, P, i, i64, i64, i64, P, P, P]
        L.bem_dense_matvec.restype = i
        L.bem_gmres_cgs2.argtypes = [P, i64, i, i64, P, P, P, ctypes.c_double, P]
        L.bem_gmres_cgs2.restype = i
        L.bem_regular_assemble_sl_dl.argtypes = [i, SP, SP, BP, BP, i, QP, d, d, i, d, i, P, P, i64, i, i, P]
        L.bem_regular_assemble_sl_dl.restype = i
        L.bem_regular_assemble_sl_dl_symmetric.argtypes = [i, SP, BP, BP, i, P, i64, QP, d, d, i, d, i, P, P, i64, i, P]
        L.bem_regular_assemble_sl_dl_symmetric.restype = i
        L.bem_dense_add_transpose.argtypes = [P, i, i64, i64, P]
        L.bem_dense_add_transpose.restype = i
        L.bem_dense_zero.argtypes = [P, i, i64, i64, P]
        L.bem_dense_zero.restype = i
        L.bem_dense_mirror_upper.argtypes = [P, i, i64, i64, P]
        L.bem_dense_mirror_upper.restype = i
        L.bem_regular_assemble_symmetric.argtypes = [i, SP, QP, d, d, i, d, i, P, i64, i, P]
        L.bem_regular_assemble_symmetric.restype = i
        L.bem_near_assemble_sl_dl.argtypes = [i, SP, SP, i, P, i, QP, d, d, d, P, P, i64, i, P]
        L.bem_near_assemble_sl_dl.restype = i
        L.bem_near_assemble_sl_dl_symmetric.argtypes = [i, SP, i, P, QP, d, d, d, P, P, i64, i, P]
        L.bem_near_assemble_sl_dl_symmetric.restype = i
        L.bem_singular_assemble_sl_dl.argtypes = [i, SP, SP, i, P, P, P, P, d, d, P, P, i64, i, P]
        L.bem_singular_assemble_sl_dl.restype = i
        _LIB = L
    return _LIB


def check(rc):
    """Raise the reference's exception class for a non-zero return code."""
    if rc != 0:
        msg = lib().bem_last_error().decode()
        raise _ERRORS.get(rc, RuntimeError)(f"libbempp_hip error {rc}: {msg}")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("device_interface='hip' needs a ROCm GPU; none is visible")
    return torch


def device():
    return _torch().device("cuda", _torch().cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def quad_rule(points, weights):
    r = BemQuadRule()
    n = len(weights)
    if n > QUAD_MAX_POINTS:
        raise ValueError(f"quadrature rules with more than {QUAD_MAX_POINTS} points are not supported")
    r.n_points = n
    for k in range(n):
        r.x[k] = float(points[0, k])
        r.y[k] = float(points[1, k])
        r.w[k] = float(weights[k])
    return r


# ------------------------------------------------------------------ device caches
_GRID_CACHE = {}
_SPACE_CACHE = {}


def geometry_table(grid):
    """(F, 16) float64 host table: v0, v1-v0, v2-v0, normal, integration element, diameter."""
    F = grid.number_of_elements
    table = np.zeros((F, GEOM_STRIDE), dtype=np.float64)
    v = grid.vertices
    el = grid.elements
    table[:, 0:3] = v[:, el[0]].T
    table[:, 3:6] = grid.jacobians[:, :, 0]
    table[:, 6:9] = grid.jacobians[:, :, 1]
    table[:, 9:12] = grid.normals
    table[:, 12] = grid.integration_elements
    table[:, 13] = grid.diameters
    return table


class DeviceGrid(object):
    """Geometry table and vertex ids of a grid in HBM."""

    def __init__(self, grid):
        torch = _torch()
        dev = device()
        self.geometry = torch.from_numpy(geometry_table(grid)).to(dev)
        vids = np.zeros((grid.number_of_elements, 4), dtype=np.int32)
        vids[:, :3] = grid.elements.T
        self.vertex_ids = torch.from_numpy(vids).to(dev)


def device_grid(grid):
    key = grid.id
    if key not in _GRID_CACHE:
        _GRID_CACHE[key] = DeviceGrid(grid)
    return _GRID_CACHE[key]


class DeviceSpace(object):
    """Dof map of a space in HBM plus its C descriptor."""

    def __init__(self, space, elements=None):
        torch = _torch()
        dev = device()
        g = device_grid(space.grid)
        if elements is None:
            elements = space.support_elements
        self.grid = g
        self.elements = torch.from_numpy(np.ascontiguousarray(elements, dtype=np.int32)).to(dev)
        self.local2global = torch.from_numpy(np.ascontiguousarray(space.local2global, dtype=np.int32)).to(dev)
        self.multipliers = torch.from_numpy(np.ascontiguousarray(space.local_multipliers, dtype=np.float64)).to(dev)
        self.normal_multipliers = torch.from_numpy(np.ascontiguousarray(space.normal_multipliers, dtype=np.int32)).to(dev)
        self.shapeset = SHAPESETS[space.shapeset.identifier]
        self.struct = BemSpace(
            int(len(elements)),
            self.shapeset,
            self.elements.data_ptr(),
            g.geometry.data_ptr(),
            g.vertex_ids.data_ptr(),
            self.local2global.data_ptr(),
            self.multipliers.data_ptr(),
            self.normal_multipliers.data_ptr(),
        )


def device_space(space):
    key = space.id
    if key not in _SPACE_CACHE:
        _SPACE_CACHE[key] = DeviceSpace(space)
    return _SPACE_CACHE[key]


def clear_caches():
    _GRID_CACHE.clear()
    _SPACE_CACHE.clear()


# ------------------------------------------------------------------ entry points
def fill_zero(tensor):
    check(lib().bem_fill_zero(ptr(tensor), tensor.numel() * tensor.element_size(), stream_handle()))


def _dtype_code(t):
    return DTYPES[str(t.dtype).replace("torch.", "")]


def regular_assemble(op, test, trial, grids_identical, rule, wavenumber, precision, out, accumulate=0,
                     near_factor=DEFAULT_NEAR_FIELD_FACTOR, kahan=False):
    """Regular pairs of localised spaces (DP0/DP1): direct stores into out."""
    k = complex(wavenumber)
    check(
        lib().bem_regular_assemble(
            OPERATORS[op], ctypes.byref(test.struct), ctypes.byref(trial.struct), int(grids_identical),
            ctypes.byref(rule), k.real, k.imag, precision, float(near_factor), int(bool(kahan)), ptr(out),
            out.shape[-1], _dtype_code(out), int(accumulate), stream_handle(),
        )
    )


def regular_assemble_blocked(op, test_blocks, trial_blocks, grids_identical, rule, wavenumber, precision, out,
                             near_factor=DEFAULT_NEAR_FIELD_FACTOR, kahan=False):
    """Regular pairs of spaces with shared dofs (P1, RWG/SNC): LDS-tiled sums added to out."""
    k = complex(wavenumber)
    check(
        lib().bem_regular_assemble_blocked(
            OPERATORS[op], ctypes.byref(test_blocks.space_struct), ctypes.byref(trial_blocks.space_struct),
            ctypes.byref(test_blocks.struct), ctypes.byref(trial_blocks.struct), int(grids_identical),
            ctypes.byref(rule), k.real, k.imag, precision, float(near_factor), int(bool(kahan)), ptr(out),
            out.shape[-1], _dtype_code(out), stream_handle(),
        )
    )


def near_assemble(op, test, trial, pairs, grids_identical, rule, wavenumber, near_factor, out):
    """FP64 near-field pairs of the mixed mode (pairs: (n, 2) int32 device tensor of candidates)."""
    k = complex(wavenumber)
    check(
        lib().bem_near_assemble(
            OPERATORS[op], ctypes.byref(test.struct), ctypes.byref(trial.struct), int(pairs.shape[0]), ptr(pairs),
            int(grids_identical), ctypes.byref(rule), k.real, k.imag, float(near_factor), ptr(out), out.shape[-1],
            _dtype_code(out), stream_handle(),
        )
    )


def singular_assemble(op, test, trial, pairs, test_points, trial_points, weights, wavenumber, out):
    k = complex(wavenumber)
    n = 0 if pairs is None else pairs.shape[0]
    check(
        lib().bem_singular_assemble(
            OPERATORS[op], ctypes.byref(test.struct), ctypes.byref(trial.struct), n, ptr(pairs), ptr(test_points),
            ptr(trial_points), ptr(weights), k.real, k.imag, ptr(out), out.shape[-1],
            DTYPES[str(out.dtype).replace("torch.", "")], stream_handle(),
        )
    )


def potential_evaluate(op, space, coefficients, points, rule, wavenumber, precision, out):
    """coefficients: complex128 (F * nshape) device tensor; points: (M, 3) float64 device tensor."""
    torch = _torch()
    k = complex(wavenumber)
    L = lib()
    n = points.shape[0]
    need = L.bem_potential_workspace_size(OPERATORS[op], ctypes.byref(space.struct), ctypes.byref(rule), n, precision)
    ws = torch.empty(need, dtype=torch.uint8, device=points.device)
    check(
        L.bem_potential_evaluate(
            OPERATORS[op], ctypes.byref(space.struct), ptr(coefficients), n, ptr(points), ctypes.byref(rule), k.real,
            k.imag, precision, ptr(out), DTYPES[str(out.dtype).replace("torch.", "")], ptr(ws), need, stream_handle(),
        )
    )
    return ws


def kbatch_assemble(test, trial, grids_identical, rule, wavenumbers, precision, near_factor, config, out,
                    near_pairs=None, singular=None, parts=KBATCH_ALL):
    """Helmholtz single layer (DP0 x DP0) for many real wavenumbers: out is (nk, rows, cols).

    config: (test_tile, workgroups_per_cu). singular: DeviceSingularRule or None.
    """
    ks = np.ascontiguousarray(np.asarray(wavenumbers, dtype=np.float64))
    nk = ks.shape[0]
    if out.dim() != 3 or out.shape[0] != nk:
        raise ValueError("out must have shape (number of wavenumbers, rows, cols)")
    cfg = BemKBatchConfig(int(config[0]), int(config[1]))
    n_near = 0 if near_pairs is None else int(near_pairs.shape[0])
    if singular is None:
        n_sing, sp, tp, spp, w = 0, None, None, None, None
    else:
        n_sing, sp, tp, spp, w = int(singular.pairs.shape[0]), singular.pairs, singular.test_points, singular.trial_points, singular.weights
    check(
        lib().bem_kbatch_assemble(
            ctypes.byref(test.struct), ctypes.byref(trial.struct), int(grids_identical), ctypes.byref(rule), nk,
            ks.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), precision, int(parts), float(near_factor),
            ctypes.byref(cfg),
            ptr(out), out.shape[-1], out.shape[-1] * out.shape[-2], _dtype_code(out), n_near, ptr(near_pairs),
            n_sing, ptr(sp), ptr(tp), ptr(spp), ptr(w), stream_handle(),
        )
    )


def adjoint_single_layer(space, rule, wavenumber, precision, near_factor, x, y, grad, near_pairs=None,
                         singular=None):
    """Add the vertex gradient of Re(x^H V y) to grad ((n_vertices, 3) float64 device tensor)."""
    n_near = 0 if near_pairs is None else int(near_pairs.shape[0])
    if singular is None:
        n_sing, sp, tp, spp, w = 0, None, None, None, None
    else:
        n_sing, sp, tp, spp, w = int(singular.pairs.shape[0]), singular.pairs, singular.test_points, \
            singular.trial_points, singular.weights
    check(
        lib().bem_adjoint_single_layer(
            ctypes.byref(space.struct), ctypes.byref(rule), float(wavenumber), precision, float(near_factor), ptr(x),
            ptr(y), n_near, ptr(near_pairs), n_sing, ptr(sp), ptr(tp), ptr(spp), ptr(w), ptr(grad), stream_handle(),
        )
    )


def adjoint_potential_single_layer(space, rule, wavenumber, coefficients, points, cotangent, grad_coefficients=None,
                                   grad_vertices=None):
    """Add P^H g to grad_coefficients and the vertex gradient of Re(g^H P c) to grad_vertices."""
    check(
        lib().bem_adjoint_potential_single_layer(
            ctypes.byref(space.struct), ctypes.byref(rule), float(wavenumber), ptr(coefficients),
            int(points.shape[0]), ptr(points), ptr(cotangent), ptr(grad_coefficients), ptr(grad_vertices),
            stream_handle(),
        )
    )


def dense_matvec(A, x, trans=0, out=None):
    """op(A) x with FP64 accumulation; A: (rows, cols) device tensor, x: complex128 vector."""
    torch = _torch()
    if A.dim() != 2 or A.stride(1) != 1:
        raise ValueError("A must be a row-major 2-d tensor")
    rows, cols = A.shape
    n_out = rows if trans == 0 else cols
    n_in = cols if trans == 0 else rows
    if x.shape != (n_in,) or x.dtype != torch.complex128 or not x.is_contiguous():
        raise ValueError(f"x must be a contiguous complex128 vector of length {n_in}")
    if out is None:
        out = torch.empty(n_out, dtype=torch.complex128, device=A.device)
    check(lib().bem_dense_matvec(int(trans), ptr(A), _dtype_code(A), rows, cols, A.stride(0), ptr(x), ptr(out),
                                 stream_handle()))
    return out


def gmres_cgs2(V, k, w, out, v_next=None, eps=0.0):
    """CGS2 of w (complex128, length n, updated in place) against rows [0, k) of V ((m, n) complex128,
    row-major); out: complex128 device buffer of >= 2k + 2 values receiving c1, |w|^2, c2, |w_out|^2.
    v_next (optional complex128 vector, e.g. V[k]): receives w_out / |w_out|, or w_out when
    |w_out| <= eps |w|."""
    torch = _torch()
    n = w.shape[0]
    if V.dtype != torch.complex128 or V.dim() != 2 or V.stride(1) != 1 or V.shape[1] != n or not 1 <= k <= V.shape[0]:
        raise ValueError("V must be a row-major (m, n) complex128 tensor with k <= m rows used")
    if w.dtype != torch.complex128 or not w.is_contiguous() or out.dtype != torch.complex128 or out.numel() < 2 * k + 2:
        raise ValueError("w must be a contiguous complex128 vector and out hold 2k + 2 complex128 values")
    if v_next is not None and (v_next.dtype != torch.complex128 or not v_next.is_contiguous() or v_next.shape != (n,)):
        raise ValueError("v_next must be a contiguous complex128 vector of length n")
    check(lib().bem_gmres_cgs2(ptr(V), V.stride(0), int(k), int(n), ptr(w), ptr(out),
                               None if v_next is None else ptr(v_next), float(eps), stream_handle()))


def regular_assemble_sl_dl(op, test, trial, test_blocks, trial_blocks, grids_identical, rule, wavenumber, precision,
                           out_single, out_double, accumulate=0, near_factor=DEFAULT_NEAR_FIELD_FACTOR, kahan=False):
    """Single layer (op) and its double layer in one pass; blocks None for localised spaces.

    With blocks, the element lists of the blocks (their space_struct) replace those of test/trial.
    """
    k = complex(wavenumber)
    tb = None if test_blocks is None else ctypes.byref(test_blocks.struct)
    sb = None if trial_blocks is None else ctypes.byref(trial_blocks.struct)
    ts = test.struct if test_blocks is None else test_blocks.space_struct
    ss = trial.struct if trial_blocks is None else trial_blocks.space_struct
    check(
        lib().bem_regular_assemble_sl_dl(
            OPERATORS[op], ctypes.byref(ts), ctypes.byref(ss), tb, sb, int(grids_identical),
            ctypes.byref(rule), k.real, k.imag, precision, float(near_factor), int(bool(kahan)), ptr(out_single),
            ptr(out_double), out_single.shape[-1], _dtype_code(out_single), int(accumulate), stream_handle(),
        )
    )


def regular_assemble_sl_dl_symmetric(op, sym, rule, wavenumber, precision, out_single, out_double, near_factor=DEFAULT_NEAR_FIELD_FACTOR,
                                     kahan=False):
    """Single layer (op) and its double layer on one P1 space, every unordered pair once.

    sym: blocking.DeviceSymmetricBlocks of the space. Both outputs are overwritten with the
    regular part (square, contiguous, same dtype). out_double None: the single layer alone.
    """
    k = complex(wavenumber)
    n = out_single.shape[0]
    for o in (out_single,) if out_double is None else (out_single, out_double):
        if o.dim() != 2 or tuple(o.shape) != (n, n) or not o.is_contiguous() or o.dtype != out_single.dtype:
            raise ValueError("outputs must be contiguous square matrices of one dtype")
    check(
        lib().bem_regular_assemble_sl_dl_symmetric(
            OPERATORS[op], ctypes.byref(sym.trial.space_struct), ctypes.byref(sym.tiles_struct),
            ctypes.byref(sym.trial.struct), int(sym.n_work), ptr(sym.work), n, ctypes.byref(rule), k.real, k.imag,
            precision, float(near_factor), int(bool(kahan)), ptr(out_single), ptr(out_double), n,
            _dtype_code(out_single), stream_handle(),
        )
    )


def regular_assemble_symmetric(op, space, rule, wavenumber, precision, out, near_factor=DEFAULT_NEAR_FIELD_FACTOR,
                               kahan=False):
    """Regular part of a symmetric single layer on one DP0 space (dof = list position): the pairs of
    the upper triangle are evaluated and each is stored with its mirror image. `out` (n x n,
    contiguous) is overwritten."""
    k = complex(wavenumber)
    n = out.shape[0]
    if out.dim() != 2 or out.shape[1] != n or not out.is_contiguous():
        raise ValueError("expected a contiguous square output")
    check(
        lib().bem_regular_assemble_symmetric(
            OPERATORS[op], ctypes.byref(space.struct), ctypes.byref(rule), k.real, k.imag, precision,
            float(near_factor), int(bool(kahan)), ptr(out), n, _dtype_code(out), stream_handle(),
        )
    )


def dense_add_transpose(a):
    """a <- a + a^T in place (square contiguous device matrix)."""
    if a.dim() != 2 or a.shape[0] != a.shape[1] or not a.is_contiguous():
        raise ValueError("expected a contiguous square matrix")
    check(lib().bem_dense_add_transpose(ptr(a), _dtype_code(a), a.shape[0], a.shape[1], stream_handle()))


def dense_zero(a):
    """a <- 0 (square contiguous device matrix) with the library's fill kernel."""
    if a.dim() != 2 or a.shape[0] != a.shape[1] or not a.is_contiguous():
        raise ValueError("expected a contiguous square matrix")
    check(lib().bem_dense_zero(ptr(a), _dtype_code(a), a.shape[0], a.shape[1], stream_handle()))


def dense_mirror_upper(a):
    """Lower triangle <- upper triangle, zero diagonal (square contiguous device matrix)."""
    if a.dim() != 2 or a.shape[0] != a.shape[1] or not a.is_contiguous():
        raise ValueError("expected a contiguous square matrix")
    check(lib().bem_dense_mirror_upper(ptr(a), _dtype_code(a), a.shape[0], a.shape[1], stream_handle()))


def near_assemble_sl_dl(op, test, trial, pairs, grids_identical, rule, wavenumber, near_factor, out_single,
                        out_double):
    k = complex(wavenumber)
    check(
        lib().bem_near_assemble_sl_dl(
            OPERATORS[op], ctypes.byref(test.struct), ctypes.byref(trial.struct), int(pairs.shape[0]), ptr(pairs),
            int(grids_identical), ctypes.byref(rule), k.real, k.imag, float(near_factor), ptr(out_single),
            ptr(out_double), out_single.shape[-1], _dtype_code(out_single), stream_handle(),
        )
    )


def near_assemble_sl_dl_symmetric(op, space, pairs, rule, wavenumber, near_factor, out_single, out_double):
    """Near pairs of the symmetric pass: each (e, f) of `pairs` once, both orientations added."""
    k = complex(wavenumber)
    check(
        lib().bem_near_assemble_sl_dl_symmetric(
            OPERATORS[op], ctypes.byref(space.struct), int(pairs.shape[0]), ptr(pairs), ctypes.byref(rule), k.real,
            k.imag, float(near_factor), ptr(out_single), ptr(out_double), out_single.shape[-1],
            _dtype_code(out_single), stream_handle(),
        )
    )


def singular_assemble_sl_dl(op, test, trial, pairs, test_points, trial_points, weights, wavenumber, out_single,
                            out_double):
    k = complex(wavenumber)
    n = 0 if pairs is None else pairs.shape[0]
    check(
        lib().bem_singular_assemble_sl_dl(
            OPERATORS[op], ctypes.byref(test.struct), ctypes.byref(trial.struct), n, ptr(pairs), ptr(test_points),
            ptr(trial_points), ptr(weights), k.real, k.imag, ptr(out_single), ptr(out_double), out_single.shape[-1],
            _dtype_code(out_single), stream_handle(),
        )
    )
