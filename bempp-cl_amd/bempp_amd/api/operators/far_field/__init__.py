"""Far-field operators (bempp_cl/api/operators/far_field): Helmholtz and Maxwell."""

from . import helmholtz, maxwell  # noqa: F401
