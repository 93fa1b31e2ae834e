"""The reference's backend interface, implemented on libbempp_hip.so.

bempp-cl dispatches dense assembly to a backend module per device interface
(bempp_cl/core/dispatcher.py:4-55) whose three functions have these signatures
(bempp_cl/core/numba_assemblers.py:6-164, bempp_cl/core/opencl_assemblers.py):

* dense_assembler(device_interface, operator_descriptor, domain, dual_to_range, parameters,
  result): adds the regular part to the caller's zero-initialised numpy `result`
  (rows = dual_to_range dofs, cols = domain dofs); called by assemble_dense
  (bempp_cl/core/dense_assembler.py:49-100);
* singular_assembler(device_interface, operator_descriptor, grid, domain, dual_to_range,
  test_points, trial_points, quad_weights, test_elements, trial_elements, test_offsets,
  trial_offsets, weights_offsets, number_of_quad_points, kernel_options, result): fills the flat
  `result[p * nt * ns + i * ns + j]` with the local block of singular pair p (edge lengths and
  normal multipliers in, the local multipliers left to the caller's scatter); called by
  assemble_singular_part (bempp_cl/core/singular_assembler.py:69-140) on the localised spaces;
* potential_assembler(device_interface, space, operator_descriptor, points, parameters) ->
  evaluator(x): x holds one coefficient per (grid element, shape function), multipliers applied;
  the evaluator returns the kernel_dimension * n_points values point by point (the caller
  reshapes them with order="F", bempp_cl/core/dense_potential_assembler.py:26-36).

A bempp-cl checkout gains device_interface="hip" with a module bempp_cl/core/hip_assemblers.py
that re-exports these functions and one "hip" branch per dispatcher function.

The reference's objects are read only through attributes they expose: the descriptor's
kernel_type, assembly_type, options, precision, is_complex and kernel_dimension; a space's grid
(compared with ==), local2global, local_multipliers, normal_multipliers, support_elements,
number_of_shape_functions, global_dof_count and shapeset.identifier; the arrays of
grid.data("double") (vertices, elements, jacobians, normals, integration_elements, diameters,
centroids); parameters.quadrature.regular. Everything is computed by the HIP library; device
copies made for a call are released when it returns.
"""

import itertools as _itertools
import weakref as _weakref

import numpy as _np

_VIEW_IDS = _itertools.count()
_SCALAR_SHAPESETS = ("p0_discontinuous", "p1_discontinuous")
_PRECISION_TYPES = {"single": ("float32", "complex64"), "double": ("float64", "complex128")}


def _check_interface(device_interface):
    if str(device_interface).split("_")[0] != "hip":
        raise ValueError(f"hip_assemblers serves device_interface 'hip', got {device_interface!r}")


class _GridView(object):
    """The geometry of a reference grid, from the arrays of grid.data("double")."""

    def __init__(self, grid):
        data = grid.data("double")
        self.vertices = _np.ascontiguousarray(data.vertices, dtype=_np.float64)
        self.elements = _np.ascontiguousarray(data.elements).astype(_np.int64)
        self.jacobians = _np.asarray(data.jacobians, dtype=_np.float64)
        self.normals = _np.asarray(data.normals, dtype=_np.float64)
        self.integration_elements = _np.asarray(data.integration_elements, dtype=_np.float64)
        self.diameters = _np.asarray(data.diameters, dtype=_np.float64)
        self.centroids = _np.asarray(data.centroids, dtype=_np.float64)
        self.number_of_elements = self.elements.shape[1]
        self.number_of_vertices = self.vertices.shape[1]
        self.id = f"hipview-grid{next(_VIEW_IDS)}"
        self.topology_id = self.id


class _Shapeset(object):
    def __init__(self, identifier):
        self.identifier = identifier


class _SpaceView(object):
    """The dof map of a reference space on a _GridView, with the attributes the HIP host code reads."""

    def __init__(self, space, grid_view):
        self.grid = grid_view
        self.local2global = _np.ascontiguousarray(space.local2global, dtype=_np.int32)
        self.local_multipliers = _np.ascontiguousarray(space.local_multipliers, dtype=_np.float64)
        self.normal_multipliers = _np.ascontiguousarray(space.normal_multipliers, dtype=_np.int32)
        self.support_elements = _np.ascontiguousarray(space.support_elements, dtype=_np.int64)
        self.number_of_support_elements = len(self.support_elements)
        self.support = _np.zeros(grid_view.number_of_elements, dtype=bool)
        self.support[self.support_elements] = True
        self.number_of_shape_functions = int(space.number_of_shape_functions)
        self.global_dof_count = int(space.global_dof_count)
        self.shapeset = _Shapeset(space.shapeset.identifier)
        self.codomain_dimension = 1 if self.shapeset.identifier in _SCALAR_SHAPESETS else 3
        # direct stores need every (support element, shape function) to own its dof, multiplier 1
        dofs = self.local2global[self.support_elements].ravel()
        mults = self.local_multipliers[self.support_elements].ravel()
        self.is_localised = bool(len(_np.unique(dofs)) == len(dofs) and _np.all(mults == 1.0))
        self.id = f"hipview-space{next(_VIEW_IDS)}"


def _views(*spaces):
    """Views of reference spaces. Spaces on equal grids share one grid view, so the views' grids
    compare equal; one space passed twice gets one view (the symmetric traversals need it)."""
    grids, made, views = [], [], []
    for s in spaces:
        for other, v in made:
            if s is other:
                views.append(v)
                break
        else:
            for g, gv in grids:
                if s.grid == g:
                    break
            else:
                gv = _GridView(s.grid)
                grids.append((s.grid, gv))
            v = _SpaceView(s, gv)
            made.append((s, v))
            views.append(v)
    return views


def _release(views):
    from bempp_amd.core import release

    for v in views:
        release(v.grid, (v,))


def dense_assembler(device_interface, operator_descriptor, domain, dual_to_range, parameters, result):
    """Add the regular part of the operator to `result` (numpy, C-contiguous, (rows, cols))."""
    import torch
    from bempp_amd.core.dense_assembler import hip_dense_assemble

    _check_interface(device_interface)
    dom, dual = _views(domain, dual_to_range)
    rows, cols = dual.global_dof_count, dom.global_dof_count
    if result.shape != (rows, cols):
        raise ValueError(f"result must have shape ({rows}, {cols})")
    try:
        mat = hip_dense_assemble(dom, dual, parameters, operator_descriptor, singular=False)
        torch.cuda.synchronize()
        result += mat.cpu().numpy().astype(result.dtype, copy=False)
    finally:
        _release((dom,) if dom is dual else (dom, dual))


def singular_assembler(device_interface, operator_descriptor, grid, domain, dual_to_range, test_points, trial_points,
                       quad_weights, test_elements, trial_elements, test_offsets, trial_offsets, weights_offsets,
                       number_of_quad_points, kernel_options, result):
    """Fill `result` (flat, nt * ns values per pair) with the local blocks of the singular pairs.

    The library's singular kernel adds each block at (local2global[e, i], local2global[f, j])
    times the local multipliers. Pair p gets a test element and a trial element of its own whose
    dof maps point at rows p * nt + i and columns j of an (n_pairs * nt, ns) matrix with unit
    multipliers, so the matrix in row-major order is the reference's flat layout.
    """
    import torch
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import _kernel_name, _wavenumber

    _check_interface(device_interface)
    op = _kernel_name(operator_descriptor)
    if op not in hb.OPERATORS:
        raise KeyError(f"Unsupported kernel {op}")
    gv = _GridView(grid)
    te = _np.asarray(test_elements, dtype=_np.int64)
    se = _np.asarray(trial_elements, dtype=_np.int64)
    n = len(te)
    nt, ns = int(dual_to_range.number_of_shape_functions), int(domain.number_of_shape_functions)
    if result.shape != (n * nt * ns,):
        raise ValueError(f"result must be flat with {n * nt * ns} values")
    if n == 0:
        return
    dev = hb.device()
    table = hb.geometry_table(gv)
    keep = []

    def side(elements, nshape, l2g, normal_multipliers, identifier):
        t = {
            "elements": torch.arange(n, dtype=torch.int32, device=dev),
            "geometry": torch.from_numpy(_np.ascontiguousarray(table[elements])).to(dev),
            "vertex_ids": torch.zeros((n, 4), dtype=torch.int32, device=dev),
            "local2global": torch.from_numpy(_np.ascontiguousarray(l2g, dtype=_np.int32)).to(dev),
            "multipliers": torch.ones((n, nshape), dtype=torch.float64, device=dev),
            "normal_multipliers": torch.from_numpy(
                _np.ascontiguousarray(_np.asarray(normal_multipliers)[elements], dtype=_np.int32)).to(dev),
        }
        keep.append(t)
        s = hb.BemSpace(n, hb.SHAPESETS[identifier], *(t[k].data_ptr() for k in (
            "elements", "geometry", "vertex_ids", "local2global", "multipliers", "normal_multipliers")))
        return _StructHolder(s)

    test = side(te, nt, _np.arange(n * nt).reshape(n, nt), dual_to_range.normal_multipliers,
                dual_to_range.shapeset.identifier)
    trial = side(se, ns, _np.tile(_np.arange(ns), (n, 1)), domain.normal_multipliers, domain.shapeset.identifier)
    pairs = _np.zeros((n, 6), dtype=_np.int32)
    pairs[:, 0] = pairs[:, 1] = _np.arange(n)
    pairs[:, 2] = test_offsets
    pairs[:, 3] = trial_offsets
    pairs[:, 4] = weights_offsets
    pairs[:, 5] = number_of_quad_points
    tp = torch.from_numpy(_np.ascontiguousarray(_np.asarray(test_points, dtype=_np.float64).T)).to(dev)
    sp = torch.from_numpy(_np.ascontiguousarray(_np.asarray(trial_points, dtype=_np.float64).T)).to(dev)
    w = torch.from_numpy(_np.ascontiguousarray(quad_weights, dtype=_np.float64)).to(dev)
    out = torch.zeros((n * nt, ns), dtype=getattr(torch, _np.dtype(result.dtype).name), device=dev)
    hb.singular_assemble(op, test, trial, torch.from_numpy(pairs).to(dev), tp, sp, w,
                         _wavenumber(operator_descriptor), out)
    torch.cuda.synchronize()
    result[:] = out.cpu().numpy().ravel()


class _StructHolder(object):
    """Carries a BemSpace as the .struct the hip_backend wrappers expect."""

    def __init__(self, struct):
        self.struct = struct


def potential_assembler(device_interface, space, operator_descriptor, points, parameters):
    """Return evaluator(x) -> kernel_dimension * n_points values, point by point."""
    import torch
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.core.dense_assembler import _kernel_name, _wavenumber

    _check_interface(device_interface)
    op = _kernel_name(operator_descriptor)
    if op not in hb.OPERATORS:
        raise KeyError(f"Unsupported kernel {op}")
    (view,) = _views(space)
    dev = hb.device()
    pts = _np.asarray(points, dtype=_np.float64)
    if pts.ndim != 2 or pts.shape[0] != 3:
        raise ValueError("points must have shape (3, n_points)")
    dpts = torch.from_numpy(_np.ascontiguousarray(pts.T)).to(dev)
    qp, qw = triangle_rule(parameters.quadrature.regular)
    rule = hb.quad_rule(qp, qw)
    precision = operator_descriptor.precision
    mode = hb.PRECISION_MIXED if precision == "single" else hb.PRECISION_FP64
    real_t, complex_t = _PRECISION_TYPES[precision]
    out_t = getattr(torch, complex_t)
    kdim = int(operator_descriptor.kernel_dimension)
    k = _wavenumber(operator_descriptor)
    dspace = hb.device_space(view)

    def evaluator(x):
        coeffs = torch.from_numpy(_np.ascontiguousarray(_np.asarray(x, dtype=_np.complex128))).to(dev)
        out = torch.empty((kdim, pts.shape[1]), dtype=out_t, device=dev)
        ws = hb.potential_evaluate(op, dspace, coeffs, dpts, rule, k, mode, out)
        torch.cuda.synchronize()
        del ws
        res = out.cpu().numpy().T.ravel()
        return res if operator_descriptor.is_complex else res.real.astype(real_t)

    # the device copies of the space live as long as the evaluator
    _weakref.finalize(evaluator, _release, (view,))
    return evaluator
