"""One-dimensional Gauss-Legendre rules on [0, 1].

Counterpart of bempp_cl/api/integration/gauss.py:1042-1055 (`rule`). The reference
reads nodes/weights on [-1, 1] from a 15-digit table, listed in descending order, and
maps them by x -> (1 + x) / 2, w -> w / 2. Here the nodes come from
numpy.polynomial.legendre.leggauss (full double precision); they are listed in the
same descending order so that tensor rules built from them enumerate points in the
reference's order.
"""

import numpy as _np


def rule(order):
    """Return (points, weights) of the `order`-point Gauss rule on [0, 1] (1 <= order <= 30)."""
    if order < 1 or order > 30:
        raise ValueError(f"Gauss quadrature order must be between 1 and 30. Provided: {order}")
    x, w = _np.polynomial.legendre.leggauss(int(order))
    x = x[::-1].copy()
    w = w[::-1].copy()
    return 0.5 * (1.0 + x), 0.5 * w
