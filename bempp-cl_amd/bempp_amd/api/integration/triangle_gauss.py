"""Symmetric Gauss rules on the reference triangle (0,0), (1,0), (0,1).

Counterpart of bempp_cl/api/integration/triangle_gauss.py:2928-2940 (`rule`).
The reference stores the HyENA tables of symmetric rules; this module builds the
same low-order symmetric rules from their closed forms instead of a table:

* order 1: centroid rule;
* order 2: 3-point rule with barycentric orbit (2/3, 1/6, 1/6);
* order 3: 4-point rule (centroid with weight -27/48 plus the orbit (3/5, 1/5, 1/5));
* order 4: 6-point rule of Strang and Fix / Dunavant (two 3-point orbits);
* order 5: 7-point Radon rule (centroid plus two 3-point orbits, closed form in sqrt(15)).

Orders 6 to 20 (12 to 79 points) have no closed form; they are served from
triangle_rules.npz, the reference's table of symmetric rules kept as data (same points,
same order, triangle_gauss.py:143-173, 175-2925), written by make_triangle_table.py.

The point order inside each rule follows the reference (triangle_gauss.py:2935-2939:
points are the 2nd and 3rd barycentric coordinate of each table row), so that sums
over quadrature points run in the same order. Weights sum to 1/2, the area of the
reference triangle.
"""

import os as _os

import numpy as _np

_TABLE_PATH = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "triangle_rules.npz")
_TABLE = {}
HIGHEST_ORDER = 20


def _table_rule(order):
    if order not in _TABLE:
        with _np.load(_TABLE_PATH) as t:
            _TABLE[order] = (_np.asfortranarray(t[f"p{order}"]), _np.array(t[f"w{order}"]))
    p, w = _TABLE[order]
    return p.copy(order="F"), w.copy()

# Orbit parameters of the order-4 rule (Dunavant 1985, degree 4, 6 points), given to
# 20 significant digits; exactness for all monomials of degree <= 4 is checked in tests.
_D4_A1 = 0.44594849091596488632
_D4_W1 = 0.22338158967801146570
_D4_A2 = 0.091576213509770743460
_D4_W2 = 0.10995174365532186764


def _s21(a):
    """The 3 barycentric permutations (1-2a, a, a), (a, 1-2a, a), (a, a, 1-2a)."""
    b = 1.0 - 2.0 * a
    return [(b, a, a), (a, b, a), (a, a, b)]


def _barycentric_rule(order):
    """Return (list of barycentric triples, list of weights summing to 1)."""
    if order == 1:
        return [(1.0 / 3, 1.0 / 3, 1.0 / 3)], [1.0]
    if order == 2:
        a = 1.0 / 6
        return [(1 - 2 * a, a, a), (a, 1 - 2 * a, a), (a, a, 1 - 2 * a)], [1.0 / 3] * 3
    if order == 3:
        pts = [(1.0 / 3, 1.0 / 3, 1.0 / 3)] + _s21(0.2)
        return pts, [-27.0 / 48] + [25.0 / 48] * 3
    if order == 4:
        o1 = _s21(_D4_A1)
        o2 = _s21(_D4_A2)
        # Row order of the reference table: o1[0], o2[0], o1[1], o1[2], o2[1], o2[2]
        pts = [o1[0], o2[0], o1[1], o1[2], o2[1], o2[2]]
        wts = [_D4_W1, _D4_W2, _D4_W1, _D4_W1, _D4_W2, _D4_W2]
        return pts, wts
    if order == 5:
        s15 = _np.sqrt(15.0)
        a1 = (6.0 - s15) / 21.0
        a2 = (6.0 + s15) / 21.0
        w1 = (155.0 - s15) / 1200.0
        w2 = (155.0 + s15) / 1200.0
        o1 = _s21(a2)
        o2 = _s21(a1)
        pts = [(1.0 / 3, 1.0 / 3, 1.0 / 3), o1[0], o2[0], o1[1], o1[2], o2[1], o2[2]]
        wts = [9.0 / 40, w2, w1, w2, w2, w1, w1]
        return pts, wts
    raise ValueError(f"Symmetric Gauss quadrature order must be between 1 and 20. Provided: {order}")


SUPPORTED_ORDERS = tuple(range(1, HIGHEST_ORDER + 1))


def rule(order):
    """Return (points (2, N) Fortran-ordered, weights (N,)) of the symmetric rule of given order.

    Points are reference coordinates (xi, eta) = 2nd and 3rd barycentric coordinate,
    weights sum to 1/2 (triangle_gauss.py:2928-2940).
    """
    order = int(order)
    if order < 1 or order > HIGHEST_ORDER:
        raise ValueError(f"Symmetric Gauss quadrature order must be between 1 and 20. Provided: {order}")
    if order > 5:
        return _table_rule(order)
    bary, wts = _barycentric_rule(order)
    bary = _np.array(bary, dtype=_np.float64)
    points = _np.asfortranarray(_np.vstack([bary[:, 1], bary[:, 2]]))
    return points, 0.5 * _np.array(wts, dtype=_np.float64)


def get_number_of_quad_points(order):
    """Number of points of the rule of the given order (triangle_gauss.py:2943-2945)."""
    return len(rule(order)[1])
