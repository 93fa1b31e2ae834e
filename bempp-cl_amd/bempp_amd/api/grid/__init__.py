"""Grids: geometry, adjacency, import and sphere generators."""

from .grid import Grid  # noqa: F401
from .io import import_grid, export  # noqa: F401
from . import shapes  # noqa: F401
