"""Deterministic sphere meshes used by the benchmarks and tests.

The reference builds spheres with gmsh or loads octahedron-based meshes from a data file
(bempp_cl/api/shapes/shapes.py:114-138) that is not available; the workloads here use
icosahedral spheres instead:

* `icosphere(level)`: the regular icosahedron (vertices (0, +-1, +-phi) and cyclic
  permutations, scaled to the unit sphere, 20 outward-oriented faces) refined `level`
  times by `Grid.refine` (each triangle split into 4 through its edge midpoints, children
  ordered as in grid.py:456-462), with the new midpoints projected back onto the unit
  sphere after every level. F = 20 * 4**level triangles, V = 10 * 4**level + 2 vertices.
* `perturbed_icosphere(level, ...)`: radial perturbation by a seeded sum of real
  spherical harmonics.
"""

import numpy as _np

_PHI = (1.0 + 5.0**0.5) / 2.0

_ICO_VERTICES = _np.array(
    [
        [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
        [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
        [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
    ],
    dtype=_np.float64,
)

_ICO_FACES = _np.array(
    [
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ],
    dtype=_np.uint32,
)


def icosahedron():
    """The unit icosahedron as a Grid (12 vertices, 20 outward-oriented faces)."""
    from .grid import Grid

    v = _ICO_VERTICES / _np.linalg.norm(_ICO_VERTICES, axis=1)[:, None]
    return Grid(v.T, _ICO_FACES.T)


def icosphere(level, radius=1.0):
    """Icosahedral sphere refined `level` times with midpoints projected onto the sphere."""
    from .grid import Grid

    grid = icosahedron()
    for _ in range(int(level)):
        fine = grid.refine()
        v = fine.vertices / _np.linalg.norm(fine.vertices, axis=0)[None, :]
        grid = Grid(v, fine.elements)
    if radius != 1.0:
        grid = Grid(radius * grid.vertices, grid.elements)
    return grid


def real_spherical_harmonic(l, m, points):
    """Real spherical harmonic Y_lm (orthonormal on the unit sphere) at unit vectors (3, N)."""
    from scipy.special import sph_harm_y

    x, y, z = points
    theta = _np.arccos(_np.clip(z, -1.0, 1.0))
    phi = _np.arctan2(y, x)
    if m == 0:
        return _np.real(sph_harm_y(l, 0, theta, phi))
    ylm = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return _np.sqrt(2.0) * (-1) ** m * _np.real(ylm)
    return _np.sqrt(2.0) * (-1) ** m * _np.imag(ylm)


def smooth_radial_field(points, n_terms=16, max_degree=6, seed=4):
    """Seeded sum of n_terms real spherical harmonics, degrees l ~ U{0..max_degree}, m ~ U{-l..l}, a ~ N(0,1)."""
    rng = _np.random.default_rng(seed)
    field = _np.zeros(points.shape[1])
    terms = []
    for _ in range(n_terms):
        l = int(rng.integers(0, max_degree + 1))
        m = int(rng.integers(-l, l + 1))
        a = float(rng.standard_normal())
        terms.append((l, m, a))
        field += a * real_spherical_harmonic(l, m, points)
    return field, terms


def perturbed_icosphere(level, amplitude=0.05, n_terms=16, max_degree=6, seed=4):
    """Icosphere with vertices moved radially: x <- x (1 + amplitude * s(x) / max|s|)."""
    from .grid import Grid

    base = icosphere(level)
    unit = base.vertices
    s, _ = smooth_radial_field(unit, n_terms, max_degree, seed)
    scale = 1.0 + amplitude * s / _np.max(_np.abs(s))
    return Grid(unit * scale[None, :], base.elements)
