"""Function spaces: DP0, DP1, P1, RWG and SNC."""

from .space import function_space, FunctionSpace  # noqa: F401
