"""Quadrature rules: symmetric triangle rules, 1-D Gauss and Sauter-Schwab (Duffy) rules."""

from . import triangle_gauss, gauss, duffy_galerkin  # noqa: F401
