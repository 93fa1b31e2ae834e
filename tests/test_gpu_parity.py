"""HIP path vs the FP64 oracle on icosahedral spheres.

FP64 mode must agree with the oracle within 1e-12 relative Frobenius error; the mixed
(FP32 far-field / FP64 near-field and singular) mode within 1e-6 of the FP64 HIP result.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_fro(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ico2():
    from bempp_amd.api.grid.shapes import icosphere

    return icosphere(2)


CASES = [
    ("laplace", "single_layer", ("DP", 0), ("DP", 0), None),
    ("laplace", "double_layer", ("P", 1), ("P", 1), None),
    ("laplace", "adjoint_double_layer", ("P", 1), ("DP", 1), None),
    ("helmholtz", "single_layer", ("P", 1), ("P", 1), 4.0),
    ("helmholtz", "double_layer", ("P", 1), ("P", 1), 4.0),
    ("helmholtz", "adjoint_double_layer", ("P", 1), ("P", 1), 3.0 + 0.5j),
    ("helmholtz", "single_layer", ("DP", 0), ("P", 1), 2.0),
    ("laplace", "hypersingular", ("P", 1), ("P", 1), None),
    ("helmholtz", "hypersingular", ("P", 1), ("P", 1), 4.0),
    ("helmholtz", "hypersingular", ("DP", 1), ("DP", 1), 3.0 + 0.5j),
    ("helmholtz", "hypersingular", ("DP", 1), ("P", 1), 2.0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}_{c[1]}_{c[2][0]}{c[2][1]}_{c[3][0]}{c[3][1]}" for c in CASES])
def test_scalar_operators_vs_oracle(case, ico2, default_parameters, hip_available):
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from oracle import oracle as O

    mod, fn, dk, tk, k = case
    dom = function_space(ico2, *dk)
    dual = function_space(ico2, *tk)
    factory = getattr(getattr(B, mod), fn)
    args = (dom, dom, dual) if k is None else (dom, dom, dual, k)
    A64 = factory(*args, parameters=default_parameters, precision="double").weak_form().to_dense()
    ref = O.assemble(f"{mod}_{fn}", ico2.vertices, ico2.elements, dom, dual, 0.0 if k is None else k)
    assert rel_fro(A64, ref) < 1e-12
    A32 = factory(*args, parameters=default_parameters, precision="single").weak_form().to_dense()
    assert rel_fro(A32.astype(A64.dtype), A64) < 1e-6


@pytest.mark.parametrize("fam,k", [("laplace", None), ("helmholtz", 3.0 + 0.5j)])
def test_hypersingular_swapped_normals_vs_oracle(fam, k, ico2, default_parameters, hip_available):
    """Normal multipliers -1 on one segment: the curls and the normal product both carry them."""
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.grid import Grid
    from oracle import oracle as O

    upper = (ico2.centroids[:, 2] > 0.0).astype(np.uint32)
    grid = Grid(ico2.vertices, ico2.elements, domain_indices=upper)
    p1 = function_space(grid, "P", 1, swapped_normals=[1])
    assert (p1.normal_multipliers == -1).any() and (p1.normal_multipliers == 1).any()
    factory = getattr(B, fam).hypersingular
    args = (p1, p1, p1) if k is None else (p1, p1, p1, k)
    A64 = factory(*args, parameters=default_parameters, precision="double").weak_form().to_dense()
    ref = O.assemble(f"{fam}_hypersingular", grid.vertices, grid.elements, p1, p1, 0.0 if k is None else k)
    assert rel_fro(A64, ref) < 1e-12
    A32 = factory(*args, parameters=default_parameters, precision="single").weak_form().to_dense()
    assert rel_fro(A32.astype(A64.dtype), A64) < 1e-6


FAR = [("helmholtz", "single_layer", ("P", 1), 3.0), ("helmholtz", "double_layer", ("DP", 0), 3.0 + 0.5j),
       ("maxwell", "electric_field", ("RWG", 0), 2.0 + 0.3j), ("maxwell", "magnetic_field", ("RWG", 0), 2.0)]


@pytest.mark.parametrize("case", FAR, ids=[f"{c[0]}_{c[1]}" for c in FAR])
def test_far_fields_vs_oracle(case, ico2, default_parameters, hip_available):
    """200 directions: FP64 vs the oracle within 1e-12; complex64 output within 1e-6 of it."""
    import bempp_amd.api.operators.far_field as FF
    from bempp_amd.api import GridFunction, function_space
    from oracle import oracle as O

    fam, fn, sk, k = case
    space = function_space(ico2, *sk)
    rng = np.random.default_rng(21)
    c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
    pts = rng.standard_normal((3, 200))
    pts /= np.linalg.norm(pts, axis=0)
    ref = O.potential(f"{fam}_far_field_{fn}" if fam == "helmholtz" else f"maxwell_{fn.split('_')[0]}_far_field",
                      ico2.vertices, ico2.elements, space, c, pts, k)
    factory = getattr(getattr(FF, fam), fn)
    u64 = factory(space, pts, k, parameters=default_parameters, precision="double").evaluate(
        GridFunction(space, coefficients=c))
    assert rel_fro(u64, ref) < 1e-12
    u32 = factory(space, pts, k, parameters=default_parameters, precision="single").evaluate(
        GridFunction(space, coefficients=c))
    assert u32.dtype == np.complex64 and rel_fro(u32.astype(np.complex128), ref) < 1e-6


@pytest.mark.parametrize("k", [2.0, 1.5 + 0.4j])
def test_magnetic_field_vs_oracle(k, ico2, default_parameters, hip_available):
    """MFIE RWG x SNC and the magnetic field potential: FP64 vs the oracle within 1e-12; single
    precision (far pairs / sources in FP32) within 1e-6."""
    from bempp_amd.api import GridFunction, function_space
    from bempp_amd.api.operators.boundary.maxwell import magnetic_field
    from bempp_amd.api.operators.potential import maxwell as mpot
    from oracle import oracle as O

    rwg = function_space(ico2, "RWG", 0)
    snc = function_space(ico2, "SNC", 0)
    A64 = magnetic_field(rwg, rwg, snc, k, parameters=default_parameters).weak_form().to_dense()
    ref = O.assemble("maxwell_magnetic_field", ico2.vertices, ico2.elements, rwg, snc, k)
    assert rel_fro(A64, ref) < 1e-12
    A32 = magnetic_field(rwg, rwg, snc, k, parameters=default_parameters, precision="single").weak_form().to_dense()
    assert rel_fro(A32.astype(A64.dtype), A64) < 1e-6
    rng = np.random.default_rng(5)
    c = rng.standard_normal(rwg.global_dof_count) + 1j * rng.standard_normal(rwg.global_dof_count)
    pts = rng.standard_normal((3, 100))
    pts = 2.5 * pts / np.linalg.norm(pts, axis=0)
    u = mpot.magnetic_field(rwg, pts, k, parameters=default_parameters).evaluate(GridFunction(rwg, coefficients=c))
    uref = O.potential("maxwell_magnetic_field", ico2.vertices, ico2.elements, rwg, c, pts, k)
    assert rel_fro(u, uref) < 1e-12
    u32 = mpot.magnetic_field(rwg, pts, k, parameters=default_parameters, precision="single").evaluate(
        GridFunction(rwg, coefficients=c))
    assert u32.dtype == np.complex64 and rel_fro(u32.astype(np.complex128), uref) < 1e-6


@pytest.mark.parametrize("case", ["helmholtz_dl_constant", "maxwell_e_random", "maxwell_h_random"])
def test_far_field_low_frequency_mixed(case, default_parameters, hip_available):
    """k = 0.05 on icosphere(3): the far-field weights nearly cancel over the surface (sum n c dS = 0
    for a constant density; RWG divergences sum to zero per basis function), so the result is
    O(k) smaller than its terms. The mixed kernels keep the group sums of the weights in FP64 and
    only e^{i theta} - 1 times the weights in FP32; single must stay within 2e-6 of double."""
    import bempp_amd.api.operators.far_field as FF
    from bempp_amd.api import GridFunction, function_space
    from bempp_amd.api.grid.shapes import icosphere

    grid = icosphere(3)
    rng = np.random.default_rng(4)
    pts = rng.standard_normal((3, 300))
    pts /= np.linalg.norm(pts, axis=0)
    if case == "helmholtz_dl_constant":
        space = function_space(grid, "P", 1)
        c = np.ones(space.global_dof_count, dtype=np.complex128)
        factory = FF.helmholtz.double_layer
    else:
        space = function_space(grid, "RWG", 0)
        c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
        factory = FF.maxwell.electric_field if case == "maxwell_e_random" else FF.maxwell.magnetic_field
    fun = GridFunction(space, coefficients=c)
    u64 = factory(space, pts, 0.05, parameters=default_parameters, precision="double").evaluate(fun)
    u32 = factory(space, pts, 0.05, parameters=default_parameters, precision="single").evaluate(fun)
    err = rel_fro(u32.astype(np.complex128), u64)
    print(f"{case}: |u| {np.linalg.norm(u64):.3e}, single vs double {err:.2e}", flush=True)
    assert err < 2e-6


@pytest.mark.parametrize("k", [2.0, 1.5 + 0.4j])
def test_magnetic_field_regular_passes(k, hip_available):
    """The MFIE's regular pairs through the library's three paths on icosphere(3): FP64, mixed
    without dof blocks (one lane per trial element, atomic scatter) and mixed with RWG dof blocks
    (the tiled kernel's owner reduction). The mixed passes leave the near pairs to the FP64 near
    pass (bem_near_assemble over candidate pairs); with it they agree with FP64 to FP32 rounding,
    without it they miss the near pairs entirely."""
    import torch

    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb
    from bempp_amd.api.utils.parameters import DEFAULT_NEAR_FIELD_FACTOR as near
    from bempp_amd.core.blocking import device_blocks
    from bempp_amd.core.near_field import device_near_candidates

    grid = icosphere(3)
    rwg = function_space(grid, "RWG", 0)
    snc = function_space(grid, "SNC", 0)
    test, trial = hb.device_space(snc), hb.device_space(rwg)
    rule = hb.quad_rule(*triangle_rule(4))
    shape = (snc.global_dof_count, rwg.global_dof_count)
    outs = {}
    for name in ("fp64", "untiled", "tiled"):
        out = torch.zeros(shape, dtype=torch.complex128 if name == "fp64" else torch.complex64, device=hb.device())
        if name == "tiled":
            hb.regular_assemble_blocked("maxwell_magnetic_field", device_blocks(snc, "test_rwg_mixed"),
                                        device_blocks(rwg, "trial_rwg_mixed"), True, rule, k, hb.PRECISION_MIXED, out)
        else:
            mode = hb.PRECISION_FP64 if name == "fp64" else hb.PRECISION_MIXED
            hb.regular_assemble("maxwell_magnetic_field", test, trial, True, rule, k, mode, out, accumulate=1)
        if name != "fp64":
            outs[name + "_far"] = out.cpu().numpy().astype(np.complex128)
            hb.near_assemble("maxwell_magnetic_field", test, trial, device_near_candidates(snc, rwg, near), True, rule,
                             k, near, out)
        outs[name] = out.cpu().numpy().astype(np.complex128)
    ref = outs["fp64"]
    # the far passes alone leave out a visible part (the near pairs)
    assert rel_fro(outs["tiled_far"], ref) > 1e-3 and rel_fro(outs["untiled_far"], ref) > 1e-3
    err_untiled, err_tiled = rel_fro(outs["untiled"], ref), rel_fro(outs["tiled"], ref)
    print(f"k={k}: mixed vs FP64 untiled {err_untiled:.2e} tiled {err_tiled:.2e}", flush=True)
    assert err_untiled < 1e-6 and err_tiled < 1e-6
    # entrywise against FP64, relative to the largest entry
    scale = np.abs(ref).max()
    assert np.abs(outs["tiled"] - ref).max() < 1e-5 * scale


def test_maxwell_vs_oracle(ico2, default_parameters, hip_available):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary.maxwell import electric_field
    from oracle import oracle as O

    rwg = function_space(ico2, "RWG", 0)
    snc = function_space(ico2, "SNC", 0)
    A64 = electric_field(rwg, rwg, snc, 2.0, parameters=default_parameters).weak_form().to_dense()
    ref = O.assemble("maxwell_electric_field", ico2.vertices, ico2.elements, rwg, snc, 2.0)
    assert rel_fro(A64, ref) < 1e-12
    A32 = electric_field(rwg, rwg, snc, 2.0, parameters=default_parameters, precision="single").weak_form().to_dense()
    assert rel_fro(A32.astype(A64.dtype), A64) < 1e-6


CUBE_CASES = [
    ("laplace", "single_layer", ("P", 1), ("P", 1), None),
    ("helmholtz", "double_layer", ("P", 1), ("P", 1), 3.0 + 0.5j),
    ("helmholtz", "adjoint_double_layer", ("DP", 0), ("P", 1), 2.0),
    ("helmholtz", "hypersingular", ("P", 1), ("P", 1), 2.5),
    ("maxwell", "electric_field", ("RWG", 0), ("SNC", 0), 2.0 + 0.3j),
    ("maxwell", "magnetic_field", ("RWG", 0), ("SNC", 0), 1.5 + 0.4j),
]


@pytest.mark.parametrize("case", CUBE_CASES, ids=[f"{c[0]}_{c[1]}" for c in CUBE_CASES])
def test_operators_on_cube_grid_vs_oracle(case, helpers, default_parameters, hip_available):
    """The reference's cube_grid.msh (1468 triangles, right-angled edges and corners, element
    areas varying by 2.2x): FP64 vs the oracle within 1e-12 and single vs FP64 within 1e-6, so the
    near/far split, the dof blocks and the singular pair tables are exercised off the icosphere."""
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from oracle import oracle as O

    grid = helpers.load_grid("cube_grid")
    mod, fn, dk, tk, k = case
    dom = function_space(grid, *dk)
    dual = function_space(grid, *tk)
    factory = getattr(getattr(B, mod), fn)
    args = (dom, dom, dual) if k is None else (dom, dom, dual, k)
    A64 = factory(*args, parameters=default_parameters, precision="double").weak_form().to_dense()
    ref = O.assemble(f"{mod}_{fn}", grid.vertices, grid.elements, dom, dual, 0.0 if k is None else k)
    err64 = rel_fro(A64, ref)
    A32 = factory(*args, parameters=default_parameters, precision="single").weak_form().to_dense()
    err32 = rel_fro(A32.astype(A64.dtype), A64)
    print(f"{mod}.{fn}: FP64 vs oracle {err64:.2e}, single vs FP64 {err32:.2e}", flush=True)
    assert err64 < 1e-12 and err32 < 1e-6


def test_c1_laplace_dp0_level3(default_parameters, hip_available):
    """Laplace SL DP0 on the level-3 icosphere (1280 triangles), FP64: 1e-12 vs the oracle."""
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary.laplace import single_layer
    from oracle import oracle as O

    g = icosphere(3)
    sp = function_space(g, "DP", 0)
    A = single_layer(sp, sp, sp, parameters=default_parameters).weak_form().to_dense()
    ref = O.assemble("laplace_single_layer", g.vertices, g.elements, sp, sp)
    assert A.shape == (1280, 1280)
    assert rel_fro(A, ref) < 1e-12


@pytest.mark.parametrize("op,sk,k", [("laplace.single_layer", ("P", 1), None), ("helmholtz.double_layer", ("P", 1), 4.0),
                                     ("helmholtz.single_layer", ("DP", 0), 4.0 + 0.3j),
                                     ("maxwell.electric_field", ("RWG", 0), 2.0)])
def test_potentials_vs_oracle(op, sk, k, ico2, default_parameters, hip_available):
    import bempp_amd.api.operators.potential as P
    from bempp_amd.api import GridFunction, function_space
    from oracle import oracle as O

    space = function_space(ico2, *sk)
    rng = np.random.default_rng(0)
    g = rng.standard_normal((3, 300))
    pts = 2.0 * g / np.linalg.norm(g, axis=0)
    c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
    mod, fn = op.split(".")
    factory = getattr(getattr(P, mod), fn)
    args = (space, pts) if k is None else (space, pts, k)
    fun = GridFunction(space, coefficients=c)
    u64 = factory(*args, parameters=default_parameters, precision="double").evaluate(fun)
    name = "maxwell_electric_field" if mod == "maxwell" else f"{mod}_{fn}"
    ref = O.potential(name, ico2.vertices, ico2.elements, space, c, pts, 0.0 if k is None else k)
    assert rel_fro(u64, ref) < 1e-12
    u32 = factory(*args, parameters=default_parameters, precision="single").evaluate(fun)
    assert rel_fro(u32, u64) < 1e-6


@pytest.mark.parametrize("k", [2.5, 0.5 + 0.2j])
def test_small_kr_singular_series_vs_oracle(k, default_parameters, hip_available):
    """Level 3 (diameters ~0.15): k (diam_t + diam_f) <= 1 on every singular pair, so the singular
    kernel evaluates e^{ikr} from its Taylor series; FP64 must still agree with the oracle."""
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary import helmholtz
    from oracle import oracle as O

    g = icosphere(3)
    assert abs(k.real if isinstance(k, complex) else k) * 2 * g.diameters.max() <= 1.0
    p1 = function_space(g, "P", 1)
    for fn, name in ((helmholtz.single_layer, "helmholtz_single_layer"), (helmholtz.double_layer, "helmholtz_double_layer")):
        A = fn(p1, p1, p1, k, parameters=default_parameters, precision="double").weak_form().to_dense()
        ref = O.assemble(name, g.vertices, g.elements, p1, p1, k)
        assert rel_fro(A, ref) < 1e-12


@pytest.mark.parametrize("k", [2.0, 0.5, 8.0])
@pytest.mark.parametrize("density", ["random", "smooth"])
def test_efie_potential_far_mixed(k, density, default_parameters, hip_available):
    """E-field potential at radius 10: the divergence term of RWG densities cancels in pairs
    (each basis function has zero total divergence), which the mixed kernel must not amplify.
    k = 0.5, 2 and 8 on level 3 take the short, long and per-source (half-angle sine beyond
    |k delta| = 1) series for e^{ik delta} - 1."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.potential import maxwell as mpot

    g = bempp.shapes.icosphere(3)
    rwg = bempp.function_space(g, "RWG", 0)
    n = rwg.global_dof_count
    if density == "random":
        rng = np.random.default_rng(2)
        c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    else:
        mid = 0.5 * (g.vertices[:, g.edges[0]] + g.vertices[:, g.edges[1]])
        c = np.cos(2 * mid[0]) + 1j * np.sin(mid[1])
    pts = np.random.default_rng(3).standard_normal((3, 2048))
    pts = 10.0 * pts / np.linalg.norm(pts, axis=0)
    fun = bempp.GridFunction(rwg, coefficients=c)
    u64 = mpot.electric_field(rwg, pts, k, parameters=default_parameters, precision="double").evaluate(fun)
    u32 = mpot.electric_field(rwg, pts, k, parameters=default_parameters, precision="single").evaluate(fun)
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-6


# Regular orders beyond 5 (the reference's rules of orders 1-20, 12..79 points): order 6 has 12
# points (the tiled kernels' runtime-Q path), orders 10 and 20 have 25 and 79 (the one-lane-per-pair
# FP64 kernel, direct stores for DP0/DP1 and atomic adds for P1/RWG).
HIGH_ORDER_CASES = [
    ("laplace", "single_layer", ("DP", 0), ("DP", 0), None),
    ("helmholtz", "double_layer", ("P", 1), ("P", 1), 3.0 + 0.5j),
    ("helmholtz", "adjoint_double_layer", ("DP", 1), ("P", 1), 2.0),
    ("helmholtz", "hypersingular", ("P", 1), ("P", 1), 2.5),
    ("maxwell", "electric_field", ("RWG", 0), ("SNC", 0), 2.0),
]


@pytest.mark.parametrize("order", [6, 10, 20])
@pytest.mark.parametrize("case", HIGH_ORDER_CASES,
                         ids=[f"{c[0]}_{c[1]}_{c[2][0]}{c[2][1]}_{c[3][0]}{c[3][1]}" for c in HIGH_ORDER_CASES])
def test_high_regular_orders_vs_oracle(case, order, hip_available):
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.utils.parameters import DefaultParameters
    from oracle import oracle as O

    mod, fn, dk, tk, k = case
    grid = icosphere(1)
    params = DefaultParameters()
    params.quadrature.regular = order
    params.quadrature.singular = 4
    dom = function_space(grid, *dk)
    dual = function_space(grid, *tk)
    factory = getattr(getattr(B, mod), fn)
    rng = dom if mod == "maxwell" else dom
    args = (dom, rng, dual) if k is None else (dom, rng, dual, k)
    print(f"high order case {mod}_{fn} order {order}", flush=True)
    A64 = factory(*args, parameters=params, precision="double").weak_form().to_dense()
    op_name = f"{mod}_{fn}"
    ref = O.assemble(op_name, grid.vertices, grid.elements, dom, dual, 0.0 if k is None else k, regular_order=order)
    assert rel_fro(A64, ref) < 1e-12
    A32 = factory(*args, parameters=params, precision="single").weak_form().to_dense()
    assert rel_fro(A32.astype(A64.dtype), A64) < 1e-6


@pytest.mark.parametrize("order", [10, 20])
def test_high_order_potentials_vs_oracle(order, hip_available):
    from bempp_amd.api import GridFunction, function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.potential import helmholtz as hpot
    from bempp_amd.api.operators.potential import maxwell as mpot
    from bempp_amd.api.utils.parameters import DefaultParameters
    from oracle import oracle as O

    grid = icosphere(1)
    params = DefaultParameters()
    params.quadrature.regular = order
    rng = np.random.default_rng(4)
    pts = rng.standard_normal((3, 50))
    pts = 2.5 * pts / np.linalg.norm(pts, axis=0)
    p1 = function_space(grid, "P", 1)
    rwg = function_space(grid, "RWG", 0)
    for space, factory, name, k in ((p1, hpot.double_layer, "helmholtz_double_layer", 2.0),
                                    (rwg, mpot.electric_field, "maxwell_electric_field", 1.5)):
        c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
        u = factory(space, pts, k, parameters=params, precision="double").evaluate(GridFunction(space, coefficients=c))
        ref = O.potential(name, grid.vertices, grid.elements, space, c, pts, k, order=order)
        assert rel_fro(u, ref) < 1e-12, name


@pytest.mark.parametrize("op,sk,k", [("helmholtz.single_layer", ("P", 1), 3.0), ("helmholtz.double_layer", ("P", 1), 3.0),
                                     ("maxwell.electric_field", ("RWG", 0), 3.0),
                                     ("maxwell.magnetic_field", ("RWG", 0), 3.0)])
def test_potentials_edge_point_sets(op, sk, k, ico2, default_parameters, hip_available):
    """No points, one point, and points just off the surface (radius 1.05 on the unit sphere, a
    quarter of an element size away): shapes, FP64 against the oracle and mixed against FP64."""
    import bempp_amd.api.operators.potential as P
    from bempp_amd.api import GridFunction, function_space
    from oracle import oracle as O

    space = function_space(ico2, *sk)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(space.global_dof_count) + 1j * rng.standard_normal(space.global_dof_count)
    fun = GridFunction(space, coefficients=c)
    mod, fn = op.split(".")
    factory = getattr(getattr(P, mod), fn)
    kdim = 3 if mod == "maxwell" else 1
    name = f"{mod}_{fn}"
    g = rng.standard_normal((3, 200))
    near = 1.05 * g / np.linalg.norm(g, axis=0)
    for pts in (np.zeros((3, 0)), near[:, :1], near):
        u64 = factory(space, pts, k, parameters=default_parameters, precision="double").evaluate(fun)
        u32 = factory(space, pts, k, parameters=default_parameters, precision="single").evaluate(fun)
        assert u64.shape == (kdim, pts.shape[1]) and u32.shape == (kdim, pts.shape[1])
        if pts.shape[1] == 0:
            continue
        ref = O.potential(name, ico2.vertices, ico2.elements, space, c, pts, k)
        err64, err32 = rel_fro(u64, ref), rel_fro(u32, u64)
        print(f"{op} {pts.shape[1]} points at radius 1.05: FP64 vs oracle {err64:.1e}, mixed vs FP64 {err32:.1e}",
              flush=True)
        assert err64 < 1e-12
        assert err32 < 1e-6


@pytest.mark.parametrize("radius", [1.02, 1.2])
def test_mixed_two_grids_close(radius, hip_available):
    """Test space on a sphere of the given radius, trial space on the unit sphere (level 3): no
    singular pairs, many near pairs across the two grids; mixed within 1e-6 of FP64."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz

    g1 = bempp.shapes.icosphere(3)
    g2 = g1.with_vertices(g1.vertices * radius)
    for fam, sk in (("single_layer", ("P", 1)), ("double_layer", ("P", 1)), ("adjoint_double_layer", ("DP", 0))):
        s1, s2 = bempp.function_space(g1, *sk), bempp.function_space(g2, *sk)
        f = getattr(helmholtz, fam)
        a64 = f(s2, s1, s1, 3.0, precision="double").weak_form().A
        a32 = f(s2, s1, s1, 3.0, precision="single").weak_form().A
        err = rel_fro(a32.astype(a64.dtype), a64)
        print(f"{fam} across grids at radius {radius}: mixed vs FP64 {err:.1e}", flush=True)
        assert err < 1e-6


@pytest.mark.parametrize("k", [20.0, 40.0])
def test_mixed_large_wavenumber(k, hip_available):
    """k = 20 and 40 on level 4 (about 3 and 6 elements per wavelength): phases of many turns per
    pair; mixed within 1e-6 of FP64."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz

    sp = bempp.function_space(bempp.shapes.icosphere(4), "P", 1)
    a64 = helmholtz.single_layer(sp, sp, sp, k, precision="double").weak_form().A
    a32 = helmholtz.single_layer(sp, sp, sp, k, precision="single").weak_form().A
    assert rel_fro(a32.astype(a64.dtype), a64) < 1e-6


@pytest.mark.parametrize("k", [30.0, 30.0 + 2.0j])
def test_mixed_potentials_large_wavenumber(k, hip_available):
    """k = 30 on level 3: many phase turns inside one source group; EFIE, magnetic and Helmholtz
    potentials (radius 2; double layer at 1.05) and, for real k, the four far fields: mixed within
    1e-6 of FP64."""
    import bempp_amd.api as bempp
    import bempp_amd.api.operators.far_field as FF
    import bempp_amd.api.operators.potential as P

    g = bempp.shapes.icosphere(3)
    rwg, p1 = bempp.function_space(g, "RWG", 0), bempp.function_space(g, "P", 1)
    rng = np.random.default_rng(1)
    fr = bempp.GridFunction(rwg, coefficients=rng.standard_normal(rwg.global_dof_count) + 1j * rng.standard_normal(
        rwg.global_dof_count))
    fp = bempp.GridFunction(p1, coefficients=rng.standard_normal(p1.global_dof_count) + 1j * rng.standard_normal(
        p1.global_dof_count))
    x = rng.standard_normal((3, 400))
    xh = x / np.linalg.norm(x, axis=0)
    cases = [(P.maxwell.electric_field, rwg, fr, 2 * xh), (P.maxwell.magnetic_field, rwg, fr, 2 * xh),
             (P.helmholtz.single_layer, p1, fp, 2 * xh), (P.helmholtz.double_layer, p1, fp, 1.05 * xh)]
    if np.imag(k) == 0:
        cases += [(FF.maxwell.electric_field, rwg, fr, xh), (FF.maxwell.magnetic_field, rwg, fr, xh),
                  (FF.helmholtz.single_layer, p1, fp, xh), (FF.helmholtz.double_layer, p1, fp, xh)]
    for fac, space, fun, pts in cases:
        u64 = fac(space, pts, k, precision="double").evaluate(fun)
        u32 = fac(space, pts, k, precision="single").evaluate(fun)
        assert rel_fro(u32, u64) < 1e-6, fac


def test_open_screen_vs_oracle(hip_available):
    """A flat square screen (an open surface, 512 triangles; P1 has no dofs on its boundary):
    single layers on P1 and DP0 and the hypersingular operator, FP64 against the oracle within
    1e-12 and mixed against FP64 within 1e-6."""
    import bempp_amd.api as bempp
    from bempp_amd.api.operators.boundary import helmholtz, laplace
    from oracle import oracle as O

    n = 16
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    vertices = np.vstack([X.ravel(), Y.ravel(), np.zeros(X.size)])
    tris = []
    for i in range(n):
        for j in range(n):
            a = i * (n + 1) + j
            tris += [[a, a + n + 1, a + 1], [a + 1, a + n + 1, a + n + 2]]
    g = bempp.Grid(vertices, np.array(tris).T)
    cases = [(("P", 1), "helmholtz_single_layer", helmholtz.single_layer, 3.0),
             (("DP", 0), "laplace_single_layer", laplace.single_layer, None),
             (("P", 1), "helmholtz_hypersingular", helmholtz.hypersingular, 3.0)]
    for sk, name, f, k in cases:
        sp = bempp.function_space(g, *sk)
        args = (sp, sp, sp) if k is None else (sp, sp, sp, k)
        a64 = f(*args, precision="double").weak_form().A
        a32 = f(*args, precision="single").weak_form().A
        ref = O.assemble(name, g.vertices, g.elements, sp, sp, 0.0 if k is None else k)
        assert rel_fro(a64, ref) < 1e-12, name
        assert rel_fro(a32.astype(a64.dtype), a64) < 1e-6, name
