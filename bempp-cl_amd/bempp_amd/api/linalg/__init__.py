"""Solvers on assembled operators (bempp_cl/api/linalg/__init__.py): GMRES, CG and LU on the device."""

from .direct_solvers import compute_lu_factors, lu  # noqa: F401
from .iterative_solvers import cg, gmres  # noqa: F401
