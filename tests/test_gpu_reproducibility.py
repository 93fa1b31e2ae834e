"""Run-to-run reproducibility of the device results.

Paths where each output entry has one writer (DP0 upper triangle with its mirror copy, the
potentials' fixed-order chunk reduction) must give bit-identical results on a second call.
Paths that add tile sums with atomics (P1 and RWG dof tiles, near and singular scatters) may
differ in the order of additions only: in mixed precision by a few FP32 roundings per entry, in
FP64 by a few FP64 roundings. The bounds below are those roundings, far inside the 1e-6 (mixed)
and 1e-12 (FP64) parity bounds.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    import torch

    return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()


def _twice(fn):
    import torch

    a = fn().clone()
    torch.cuda.synchronize()
    b = fn().clone()
    torch.cuda.synchronize()
    return a, b


def test_dp0_single_layer_bitwise(hip_available):
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz

    g = bempp.shapes.icosphere(3)
    dp0 = bempp.function_space(g, "DP", 0)
    for precision in ("single", "double"):
        a, b = _twice(lambda: helmholtz.single_layer(dp0, dp0, dp0, 3.0, precision=precision).weak_form().tensor)
        print(f"DP0 single layer ({precision}): max |a - b| = {(a - b).abs().max().item():.3e}", flush=True)
        assert bool((a == b).all()), precision


def test_potentials_bitwise(hip_available):
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.potential import helmholtz as hpot
    from bempp_amd.api.operators.potential import maxwell as mpot

    g = bempp.shapes.icosphere(3)
    p1 = bempp.function_space(g, "P", 1)
    rwg = bempp.function_space(g, "RWG", 0)
    rng = np.random.default_rng(5)
    pts = rng.standard_normal((3, 2000))
    pts *= 3.0 / np.linalg.norm(pts, axis=0)
    f1 = bempp.GridFunction(p1, coefficients=rng.standard_normal(p1.global_dof_count) + 0j)
    f2 = bempp.GridFunction(rwg, coefficients=rng.standard_normal(rwg.global_dof_count)
                            + 1j * rng.standard_normal(rwg.global_dof_count))
    for precision in ("single", "double"):
        u1 = hpot.single_layer(p1, pts, 2.5, precision=precision).evaluate(f1)
        u2 = hpot.single_layer(p1, pts, 2.5, precision=precision).evaluate(f1)
        assert np.array_equal(u1, u2), precision
        e1 = mpot.electric_field(rwg, pts, 2.0, precision=precision).evaluate(f2)
        e2 = mpot.electric_field(rwg, pts, 2.0, precision=precision).evaluate(f2)
        assert np.array_equal(e1, e2), precision


@pytest.mark.parametrize("precision,bound", [("single", 2e-7), ("double", 1e-15)])
def test_tiled_paths_within_rounding(precision, bound, hip_available):
    """P1 single layer (symmetric pass), P1 double layer (ordered tiles) and the EFIE (symmetric
    pass): two calls agree to the rounding of reordered additions."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz, maxwell

    g = bempp.shapes.icosphere(3)
    p1 = bempp.function_space(g, "P", 1)
    rwg = bempp.function_space(g, "RWG", 0)
    snc = bempp.function_space(g, "SNC", 0)
    cases = {
        "P1 single layer": lambda: helmholtz.single_layer(p1, p1, p1, 3.0, precision=precision).weak_form().tensor,
        "P1 double layer": lambda: helmholtz.double_layer(p1, p1, p1, 3.0, precision=precision).weak_form().tensor,
        "EFIE": lambda: maxwell.electric_field(rwg, rwg, snc, 2.0, precision=precision).weak_form().tensor,
    }
    for name, fn in cases.items():
        a, b = _twice(fn)
        d = _rel(a, b)
        print(f"{name} ({precision}): run-to-run {d:.2e}, bitwise equal {bool((a == b).all())}", flush=True)
        assert d < bound, name
