"""Parameter object (bempp_cl/api/utils/parameters.py:22-71), restricted to dense assembly."""


class _Quadrature(object):
    def __init__(self):
        self.regular = 4
        self.singular = 4


class _DenseAssembly(object):
    def __init__(self):
        self.workgroup_size_multiple = 2


class _Assembly(object):
    def __init__(self):
        self.dense = _DenseAssembly()
        self.always_promote_to_double = False
        self.discretization_type = "galerkin"


# Near-field factor of the mixed precision mode. Measured on the level-5 sphere (Helmholtz SL+DL P1,
# k = 4): relative errors against FP64 of 1.04e-7 / 1.28e-7 at factors 2.0 and 1.5 and 1.06e-7 /
# 1.29e-7 at 1.0, because the FP32 path evaluates r - r0 and the phase relative to the FP64
# centroid distance; 1.5 keeps the near-singular pairs in FP64 with a margin.
DEFAULT_NEAR_FIELD_FACTOR = 1.5


class _Hip(object):
    """Options of the HIP device interface.

    mixed_near_field_factor: in single precision, element pairs whose centroid distance is
    below this factor times the mean element diameter are integrated in FP64, the others
    in FP32 (default DEFAULT_NEAR_FIELD_FACTOR).
    kahan: Kahan-compensated accumulation of the per-pair FP32 sums over test points.
    Off by default: on the level-5 sphere (Helmholtz P1, k = 4) it changes the relative
    error against FP64 from 1.05e-7 to 1.03e-7 and costs 23% of the single-layer kernel
    time; the error is dominated by the FP32 kernel evaluation and complex64 storage.
    symmetric: when a single layer and its double layer are assembled together on one P1
    space (assemble_weak_forms), evaluate every unordered element pair once and obtain the
    reversed pair from the same distances and phase factors (the single layer's block is the
    transpose, the double layer's only needs the other normal). False visits each ordered pair.
    """

    def __init__(self):
        self.mixed_near_field_factor = DEFAULT_NEAR_FIELD_FACTOR
        self.kahan = False
        self.symmetric = True
        self.tile_config = None


class DefaultParameters(object):
    """Default parameters: regular and singular quadrature order 4."""

    def __init__(self):
        self.quadrature = _Quadrature()
        self.assembly = _Assembly()
        self.hip = _Hip()
