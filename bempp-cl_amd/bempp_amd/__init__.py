"""bempp_amd: dense Galerkin BEM assembly and dense potentials on AMD MI355X (gfx950).

Import the API as ``import bempp_amd.api as bempp``.
"""

__version__ = "0.1.0"
