"""Boundary operators: Laplace, Helmholtz, Maxwell (electric field) and the sparse identity."""

from . import laplace, helmholtz, maxwell, sparse  # noqa: F401
