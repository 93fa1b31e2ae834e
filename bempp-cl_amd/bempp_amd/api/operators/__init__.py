"""Operator descriptors and factories (bempp_cl/api/operators/__init__.py:11-19)."""

import collections as _collections

OperatorDescriptor = _collections.namedtuple(
    "OperatorDescriptor",
    "identifier options kernel_type assembly_type precision is_complex singular_part kernel_dimension",
)

from . import boundary  # noqa: E402,F401
from . import potential  # noqa: E402,F401
from . import far_field  # noqa: E402,F401
