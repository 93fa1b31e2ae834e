"""Potential operators: Laplace, Helmholtz and Maxwell (electric field)."""

from . import laplace, helmholtz, maxwell  # noqa: F401
