"""bempp_amd.api: the bempp-cl style API of the MI355X dense BEM library.

Mirrors the parts of bempp_cl.api (bempp_cl/api/__init__.py) used by dense assembly, dense
potentials and solves on the assembled operators: grids, function spaces, grid functions,
boundary (and blocked) and potential operators, the L2 identity, GMRES / CG / LU, parameters.
The only device interface is "hip" (hand-written HIP kernels in libbempp_hip.so).
"""

import logging as _logging
import time as _time

from bempp_amd.api.utils.parameters import DefaultParameters
from bempp_amd.api.grid import Grid, import_grid, export, shapes  # noqa: F401
from bempp_amd.api.space import function_space  # noqa: F401
from bempp_amd.api.assembly.grid_function import GridFunction  # noqa: F401
from bempp_amd.api.assembly.grid_function import real_callable, complex_callable, callable  # noqa: F401,A004
from bempp_amd.api.assembly.blocked_operator import BlockedOperator  # noqa: F401
from bempp_amd.api.assembly.discrete_boundary_operator import as_matrix  # noqa: F401
from bempp_amd.api import operators  # noqa: F401
from bempp_amd.api import linalg  # noqa: F401
from bempp_amd.api.linalg import lu, compute_lu_factors, gmres, cg  # noqa: F401

LOGGER = _logging.getLogger("bempp_amd")
LOGGER.addHandler(_logging.NullHandler())

GLOBAL_PARAMETERS = DefaultParameters()
DEFAULT_DEVICE_INTERFACE = "hip"
DEFAULT_PRECISION = "double"


def assign_parameters(parameters):
    """Return the global parameters if None is given (bempp_cl/api/__init__.py)."""
    return GLOBAL_PARAMETERS if parameters is None else parameters


def log(message, level="info"):
    getattr(LOGGER, level)(message)


class Timer(object):
    """Context manager measuring wall time (bempp_cl/api/__init__.py:160-186)."""

    def __init__(self, enable_log=True, message=""):
        self.enable_log = enable_log
        self.message = message
        self.start = self.end = self.interval = 0.0

    def __enter__(self):
        self.start = _time.perf_counter()
        return self

    def __exit__(self, *args):
        self.end = _time.perf_counter()
        self.interval = self.end - self.start
        if self.enable_log:
            LOGGER.debug("%s: %.6f s", self.message, self.interval)
