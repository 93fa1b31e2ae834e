"""Sauter-Schwab (Duffy-type) rules for singular Galerkin integrals over triangle pairs.

Counterpart of bempp_cl/api/integration/duffy_galerkin.py:6-285. The rules are the
published regularising transformations of Sauter and Schwab (Boundary Element
Methods, 2011, sec. 5.2): the 4-D hypercube (xi, eta1, eta2, eta3) is mapped onto
the product of the two triangles split into 6 (coincident), 5 (edge-adjacent) or 2
(vertex-adjacent) regions, with the Jacobian of each map folded into the weights.

Conventions shared with the reference so that results agree to rounding:
* tensor points run test index outer / trial index inner, with (xi, eta1) taken from
  the test tensor point and (eta2, eta3) from the trial tensor point; within a tensor
  point pair the regions are consecutive (duffy_galerkin.py:41-58, 60-115);
* the maps are written for the triangle (0,0), (1,0), (1,1) and converted to the
  reference triangle (0,0), (1,0), (0,1) by x <- x - y (duffy_galerkin.py:200-205);
* the rules assume the shared edge is (local vertex 0, local vertex 1) and the shared
  vertex is local vertex 0; `remap_points_shared_edge` / `remap_points_shared_vertex`
  move the points to the actual shared edge or vertex (duffy_galerkin.py:229-285).
"""

import numpy as _np


def number_of_quadrature_points(order, adjacency):
    """Points per pair: 6 n^4, 5 n^4, 2 n^4 (duffy_galerkin.py:6-22)."""
    factors = {"coincident": 6, "edge_adjacent": 5, "vertex_adjacent": 2}
    if adjacency not in factors:
        raise ValueError("adjacency must be one of 'coincident', 'edge_adjacent', 'vertex_adjacent'")
    return factors[adjacency] * order**4


def _coincident_regions(xsi, e1, e2, e3):
    """Six maps of the coincident rule; each returns (tx, ty, sx, sy, weight factor)."""
    e12 = e1 * e2
    e123 = e12 * e3
    one = _np.ones_like(xsi)
    return [
        (xsi, xsi * (1 - e1 + e12), xsi * (1 - e123), xsi * (1 - e1), one),
        (xsi * (1 - e123), xsi * (1 - e1), xsi, xsi * (1 - e1 + e12), one),
        (xsi, xsi * (e1 - e12 + e123), xsi * (1 - e12), xsi * (e1 - e12), one),
        (xsi * (1 - e12), xsi * (e1 - e12), xsi, xsi * (e1 - e12 + e123), one),
        (xsi * (1 - e123), xsi * (e1 - e123), xsi, xsi * (e1 - e12), one),
        (xsi, xsi * (e1 - e12), xsi * (1 - e123), xsi * (e1 - e123), one),
    ]


def _edge_regions(xsi, e1, e2, e3):
    """Five maps of the edge-adjacent rule."""
    e12 = e1 * e2
    e123 = e12 * e3
    one = _np.ones_like(xsi)
    return [
        (xsi, xsi * e1 * e3, xsi * (1 - e12), xsi * e1 * (1 - e2), one),
        (xsi, xsi * e1, xsi * (1 - e123), xsi * e12 * (1 - e3), e2),
        (xsi * (1 - e12), xsi * e1 * (1 - e2), xsi, xsi * e123, e2),
        (xsi * (1 - e123), xsi * e12 * (1 - e3), xsi, xsi * e1, e2),
        (xsi * (1 - e123), xsi * e1 * (1 - e2 * e3), xsi, xsi * e12, e2),
    ]


def _vertex_regions(xsi, e1, e2, e3):
    """Two maps of the vertex-adjacent rule."""
    one = _np.ones_like(xsi)
    return [
        (xsi, xsi * e1, xsi * e2, xsi * e2 * e3, one),
        (xsi * e2, xsi * e2 * e3, xsi, xsi * e1, one),
    ]


def rule(order, adjacency):
    """Return (test_points (2,N), trial_points (2,N), weights (N,)) of the singular rule."""
    from .gauss import rule as gauss_rule

    if adjacency not in ["coincident", "edge_adjacent", "vertex_adjacent"]:
        raise ValueError("Unknown adjacency.")

    x1d, w1d = gauss_rule(order)
    n = len(w1d)
    # 2-D tensor points: index i*n + j holds (x1d[j], x1d[i]) with weight w1d[i] w1d[j]
    tp0 = _np.tile(x1d, n)
    tp1 = _np.repeat(x1d, n)
    tw = _np.repeat(w1d, n) * _np.tile(w1d, n)
    # pairs of tensor points: test index outer, trial index inner
    xsi = _np.repeat(tp0, n * n)
    e1 = _np.repeat(tp1, n * n)
    e2 = _np.tile(tp0, n * n)
    e3 = _np.tile(tp1, n * n)
    base_w = _np.repeat(tw, n * n) * _np.tile(tw, n * n)

    if adjacency == "coincident":
        regions = _coincident_regions(xsi, e1, e2, e3)
        jac = xsi**3 * e1**2 * e2
    elif adjacency == "edge_adjacent":
        regions = _edge_regions(xsi, e1, e2, e3)
        jac = xsi**3 * e1**2
    else:
        regions = _vertex_regions(xsi, e1, e2, e3)
        jac = xsi**3 * e2

    # stack with the region index fastest
    tx = _np.stack([r[0] for r in regions], axis=1).ravel()
    ty = _np.stack([r[1] for r in regions], axis=1).ravel()
    sx = _np.stack([r[2] for r in regions], axis=1).ravel()
    sy = _np.stack([r[3] for r in regions], axis=1).ravel()
    wf = _np.stack([r[4] for r in regions], axis=1).ravel()
    weights = _np.repeat(base_w * jac, len(regions)) * wf

    points_test = _np.asfortranarray(_np.vstack([tx - ty, ty]))
    points_trial = _np.asfortranarray(_np.vstack([sx - sy, sy]))
    return points_test, points_trial, weights


def remap_points_shared_vertex(points, vertex_id):
    """Move points of a rule that meets at local vertex 0 to local vertex `vertex_id`."""
    if vertex_id == 0:
        return points
    new_points = _np.zeros_like(points)
    if vertex_id == 1:
        new_points[0, :] = 1.0 - points[0, :] - points[1, :]
        new_points[1, :] = points[1, :]
        return new_points
    if vertex_id == 2:
        new_points[0, :] = points[0, :]
        new_points[1, :] = 1.0 - points[0, :] - points[1, :]
        return new_points
    raise ValueError("vertex_id must be 0, 1 or 2")


_REF_VERTICES = _np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def remap_points_shared_edge(points, shared_vertex1, shared_vertex2):
    """Move points of a rule whose shared edge is (0, 1) to the edge (shared_vertex1, shared_vertex2).

    The affine map sends reference vertices 0, 1, 2 to the reference vertices
    shared_vertex1, shared_vertex2 and the remaining one.
    """
    v0, v1 = shared_vertex1, shared_vertex2
    v2 = 3 - v0 - v1
    origin = _REF_VERTICES[:, v0]
    amat = _np.column_stack([_REF_VERTICES[:, v1] - origin, _REF_VERTICES[:, v2] - origin])
    return _np.asfortranarray(amat @ points + origin[:, None])
