"""Symmetric single + double layer pass (bem_regular_assemble_sl_dl_symmetric).

Every unordered element pair is evaluated once: the single layer of (f, e) is the transpose
of that of (e, f) and the double layer of (f, e) comes from the same distances with the test
normal. The results must equal the pass over ordered pairs (FP64: within 1e-13) and the FP64
oracle (within 1e-12) on a grid whose elements all differ (perturbed sphere), and the mixed
symmetric pass must stay within 1e-6 of FP64.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    import torch

    return (torch.linalg.norm(a.to(b.dtype) - b) / torch.linalg.norm(b)).item()


def perturbed_sphere(level, amplitude=0.08, seed=3):
    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(level)
    rng = np.random.default_rng(seed)
    v = g.vertices * (1.0 + amplitude * rng.uniform(-1.0, 1.0, g.vertices.shape[1]))
    return g.with_vertices(v)


def fused(space, fam, k, precision, symmetric, params):
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.core.dense_assembler import hip_dense_assemble_sl_dl

    mod = getattr(B, fam)
    args = (space, space, space) if k is None else (space, space, space, k)
    sl = mod.single_layer(*args, parameters=params, precision=precision).descriptor
    dl = mod.double_layer(*args, parameters=params, precision=precision).descriptor
    params.hip.symmetric = symmetric
    try:
        return hip_dense_assemble_sl_dl(space, space, params, sl, dl)
    finally:
        params.hip.symmetric = True


CASES = [("helmholtz", 4.0), ("helmholtz", 3.0 + 0.5j), ("laplace", None)]


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}_{c[1]}" for c in CASES])
def test_symmetric_equals_ordered(case, level, default_parameters, hip_available):
    from bempp_amd.api import function_space

    fam, k = case
    p1 = function_space(perturbed_sphere(level), "P", 1)
    o1, o2 = fused(p1, fam, k, "double", False, default_parameters)
    s1, s2 = fused(p1, fam, k, "double", True, default_parameters)
    print(f"{fam} k={k} level {level}: FP64 symmetric vs ordered {rel(s1, o1):.2e} {rel(s2, o2):.2e}", flush=True)
    assert rel(s1, o1) < 1e-13
    assert rel(s2, o2) < 1e-13
    m1, m2 = fused(p1, fam, k, "single", True, default_parameters)
    print(f"  mixed symmetric vs FP64 {rel(m1, o1):.2e} {rel(m2, o2):.2e}", flush=True)
    assert rel(m1, o1) < 1e-6
    assert rel(m2, o2) < 1e-6


# k = 4: k (diam_e + diam_f) > 1 on most far pairs (the branch-free reduced phase);
# k = 1.5: below 1 on every pair (e^{ik R0} per pair times the Taylor series of e^{ik (r - R0)})
@pytest.mark.parametrize("k", [4.0, 1.5])
def test_symmetric_fp64_vs_oracle(k, default_parameters, hip_available):
    from bempp_amd.api import function_space
    from oracle import oracle as O

    grid = perturbed_sphere(3)
    p1 = function_space(grid, "P", 1)
    a1, a2 = fused(p1, "helmholtz", k, "double", True, default_parameters)
    for name, a in (("helmholtz_single_layer", a1), ("helmholtz_double_layer", a2)):
        ref = O.assemble(name, grid.vertices, grid.elements, p1, p1, k)
        err = np.linalg.norm(a.cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"{name}: symmetric FP64 vs oracle {err:.2e}", flush=True)
        assert err < 1e-12


def test_symmetric_mixed_weak_forms(default_parameters, hip_available):
    """The public API (assemble_weak_forms) takes the symmetric pass for one P1 space."""
    from bempp_amd.api import function_space
    from bempp_amd.api.assembly.boundary_operator import assemble_weak_forms
    from bempp_amd.api.operators.boundary import helmholtz

    grid = perturbed_sphere(2)
    p1 = function_space(grid, "P", 1)
    v, kd = assemble_weak_forms([
        helmholtz.single_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="single"),
        helmholtz.double_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="single"),
    ])
    rv = helmholtz.single_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="double").weak_form()
    rk = helmholtz.double_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="double").weak_form()
    for f, r in ((v, rv), (kd, rk)):
        assert np.linalg.norm(f.to_dense() - r.to_dense()) / np.linalg.norm(r.to_dense()) < 1e-6


@pytest.mark.parametrize("dtype", ["float32", "float64", "complex64", "complex128"])
@pytest.mark.parametrize("n", [1, 77, 1000])
def test_add_transpose(dtype, n, hip_available):
    import torch
    from bempp_amd.core import hip_backend as hb

    g = torch.Generator().manual_seed(n)
    a = torch.randn((n, n), generator=g, dtype=getattr(torch, dtype)).to(hb.device())
    ref = a + a.T
    hb.dense_add_transpose(a)
    torch.cuda.synchronize()
    assert torch.equal(a, ref)


@pytest.mark.parametrize("boundary", [False, True])
def test_symmetric_partial_support(boundary, default_parameters, hip_available):
    """P1 on half of a perturbed sphere (dofs on the cut excluded or kept, zero multipliers
    where a shape function has no dof): symmetric == ordered pass, FP64 vs the oracle."""
    from bempp_amd.api import function_space
    from oracle import oracle as O

    grid = perturbed_sphere(3)
    upper = np.nonzero(grid.centroids[:, 2] > 0.1)[0]
    p1 = function_space(grid, "P", 1, support_elements=upper, include_boundary_dofs=boundary)
    o1, o2 = fused(p1, "helmholtz", 3.0, "double", False, default_parameters)
    s1, s2 = fused(p1, "helmholtz", 3.0, "double", True, default_parameters)
    print(f"partial support (boundary dofs {boundary}): {rel(s1, o1):.2e} {rel(s2, o2):.2e}", flush=True)
    assert rel(s1, o1) < 1e-13 and rel(s2, o2) < 1e-13
    ref = O.assemble("helmholtz_double_layer", grid.vertices, grid.elements, p1, p1, 3.0)
    assert np.linalg.norm(s2.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-12
    m1, m2 = fused(p1, "helmholtz", 3.0, "single", True, default_parameters)
    assert rel(m1, o1) < 1e-6 and rel(m2, o2) < 1e-6


@pytest.mark.parametrize("fam,k", [("helmholtz", 3.0), ("helmholtz", 2.0 + 0.4j), ("laplace", None)])
def test_dp0_single_layer_upper_triangle(fam, k, default_parameters, hip_available):
    """DP0 single layer on one space: upper triangle + mirror copy == all ordered pairs."""
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import hip_dense_assemble, use_upper_triangle
    from oracle import oracle as O

    grid = perturbed_sphere(3)
    dp0 = function_space(grid, "DP", 0)
    assert use_upper_triangle(dp0, dp0, f"{fam}_single_layer")
    mod = getattr(B, fam)
    args = (dp0, dp0, dp0) if k is None else (dp0, dp0, dp0, k)
    res = {}
    for precision in ("double", "single"):
        d = mod.single_layer(*args, parameters=default_parameters, precision=precision).descriptor
        for sym in (False, True):
            default_parameters.hip.symmetric = sym
            try:
                res[precision, sym] = hip_dense_assemble(dp0, dp0, default_parameters, d)
            finally:
                default_parameters.hip.symmetric = True
    print(f"{fam} k={k}: upper vs ordered FP64 {rel(res['double', True], res['double', False]):.2e}", flush=True)
    assert rel(res["double", True], res["double", False]) < 1e-13
    assert rel(res["single", True], res["double", False]) < 1e-6
    ref = O.assemble(f"{fam}_single_layer", grid.vertices, grid.elements, dp0, dp0, 0.0 if k is None else k)
    assert np.linalg.norm(res["double", True].cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("dtype", ["float32", "float64", "complex64", "complex128"])
@pytest.mark.parametrize("n,ld", [(37, 37), (1000, 1000), (37, 41), (300, 320)])
def test_dense_zero_contiguous_and_strided(dtype, n, ld, hip_available):
    """bem_dense_zero: a contiguous matrix (ld == n) takes the 16-byte fill, including its tail
    bytes (37^2 float32 values are not a multiple of 16 bytes); a strided one (ld > n) the
    row kernel. The n x n block is zeroed and the padding columns keep their values."""
    import ctypes

    import torch

    from bempp_amd.core import hip_backend as hb

    torch.manual_seed(n + ld)
    a = torch.randn((n, ld), dtype=getattr(torch, dtype), device=hb.device()) + 1
    before = a.cpu()
    rc = hb.lib().bem_dense_zero(ctypes.c_void_p(a.data_ptr()), hb._dtype_code(a), n, ld, hb.stream_handle())
    assert rc == 0
    after = a.cpu()
    assert bool((after[:, :n] == 0).all())
    assert torch.equal(after[:, n:], before[:, n:])



def efie(rwg, snc, k, precision, symmetric, params):
    from bempp_amd.api.operators.boundary.maxwell import electric_field

    params.hip.symmetric = symmetric
    try:
        return electric_field(rwg, rwg, snc, k, parameters=params, precision=precision).weak_form().tensor
    finally:
        params.hip.symmetric = True


EFIE_CASES = [2.0, 3.0 + 0.5j]


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("k", EFIE_CASES)
def test_symmetric_efie_equals_ordered(k, level, default_parameters, hip_available):
    """The electric field operator on RWG x SNC over one dof map: the symmetric traversal (each
    unordered pair of the regular part once, then the transpose) against the ordered pairs."""
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import use_symmetric_efie

    grid = perturbed_sphere(level)
    rwg = function_space(grid, "RWG", 0)
    snc = function_space(grid, "SNC", 0)
    assert use_symmetric_efie(rwg, snc)
    o = efie(rwg, snc, k, "double", False, default_parameters)
    s = efie(rwg, snc, k, "double", True, default_parameters)
    print(f"EFIE k={k} level {level}: FP64 symmetric vs ordered {rel(s, o):.2e}", flush=True)
    assert rel(s, o) < 1e-13
    m = efie(rwg, snc, k, "single", True, default_parameters)
    mo = efie(rwg, snc, k, "single", False, default_parameters)
    print(f"  mixed symmetric vs FP64 {rel(m, o):.2e}, ordered mixed vs FP64 {rel(mo, o):.2e}", flush=True)
    assert rel(m, o) < 1e-6
    assert rel(mo, o) < 1e-6


def test_symmetric_efie_open_screen_vs_oracle(default_parameters, hip_available):
    """RWG x SNC on a flat screen (boundary edges carry no dofs), symmetric FP64 pass against the
    oracle within 1e-12 and the mixed symmetric pass against it within 1e-6."""
    import bempp_amd.api as bempp
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import use_symmetric_efie
    from oracle import oracle as O

    n = 10
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    vertices = np.vstack([X.ravel(), Y.ravel(), 0.05 * np.sin(3.0 * X.ravel()) * Y.ravel()])
    tris = []
    for i in range(n):
        for j in range(n):
            a = i * (n + 1) + j
            tris += [[a, a + n + 1, a + 1], [a + 1, a + n + 1, a + n + 2]]
    g = bempp.Grid(vertices, np.array(tris).T)
    rwg = function_space(g, "RWG", 0)
    snc = function_space(g, "SNC", 0)
    assert use_symmetric_efie(rwg, snc)
    k = 2.5
    ref = O.assemble("maxwell_electric_field", g.vertices, g.elements, rwg, snc, k)
    a64 = efie(rwg, snc, k, "double", True, default_parameters).cpu().numpy()
    a32 = efie(rwg, snc, k, "single", True, default_parameters).cpu().numpy()
    e64 = np.linalg.norm(a64 - ref) / np.linalg.norm(ref)
    e32 = np.linalg.norm(a32 - ref) / np.linalg.norm(ref)
    print(f"EFIE screen ({rwg.global_dof_count} dofs): symmetric FP64 {e64:.2e}, mixed {e32:.2e}", flush=True)
    assert e64 < 1e-12
    assert e32 < 1e-6


@pytest.mark.parametrize("order", [2, 6])
def test_symmetric_efie_other_rules(order, default_parameters, hip_available):
    """Regular rules other than the 6-point one (3 and 12 points: the generic point loop) in the
    symmetric EFIE pass, against the ordered pass."""
    from bempp_amd.api import function_space

    grid = perturbed_sphere(2)
    rwg = function_space(grid, "RWG", 0)
    snc = function_space(grid, "SNC", 0)
    default_parameters.quadrature.regular = order
    o = efie(rwg, snc, 2.0, "double", False, default_parameters)
    s = efie(rwg, snc, 2.0, "double", True, default_parameters)
    m = efie(rwg, snc, 2.0, "single", True, default_parameters)
    print(f"EFIE regular order {order}: FP64 symmetric vs ordered {rel(s, o):.2e}, mixed {rel(m, o):.2e}", flush=True)
    assert rel(s, o) < 1e-13
    assert rel(m, o) < 1e-6


def test_symmetric_efie_partial_support(default_parameters, hip_available):
    """RWG and SNC restricted to the same part of the sphere (edges on the part's boundary carry
    no dof): the symmetric pass against the ordered one, and mixed against FP64."""
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import use_symmetric_efie

    grid = perturbed_sphere(2)
    part = np.nonzero(grid.centroids[:, 2] > -0.3)[0]
    rwg = function_space(grid, "RWG", 0, support_elements=part)
    snc = function_space(grid, "SNC", 0, support_elements=part)
    assert use_symmetric_efie(rwg, snc)
    o = efie(rwg, snc, 2.5, "double", False, default_parameters)
    s = efie(rwg, snc, 2.5, "double", True, default_parameters)
    m = efie(rwg, snc, 2.5, "single", True, default_parameters)
    print(f"EFIE on {len(part)} of {grid.number_of_elements} triangles ({rwg.global_dof_count} dofs): "
          f"symmetric vs ordered {rel(s, o):.2e}, mixed {rel(m, o):.2e}", flush=True)
    assert rel(s, o) < 1e-13
    assert rel(m, o) < 1e-6


@pytest.mark.parametrize("level", [3, 4])
@pytest.mark.parametrize("precision_name", ["mixed", "fp64"])
def test_dp0_upper_triangle_strided_output(level, precision_name, hip_available):
    """bem_regular_assemble_symmetric into rows longer than the matrix (ld = n + 5): the n x n block
    equals the contiguous result, mirror images included, and the 5 padding columns of every row
    keep their contents. Level 3 takes the 4-element test tiles, level 4 the 16-element ones;
    complex64 output stages the mirror images in LDS, complex128 stores them directly."""
    import ctypes

    import torch

    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.integration.triangle_gauss import rule as triangle_rule
    from bempp_amd.core import hip_backend as hb

    sp = function_space(icosphere(level), "DP", 0)
    dsp = hb.device_space(sp)
    pts, w = triangle_rule(4)
    rule = hb.quad_rule(pts, w)
    precision = hb.PRECISION_MIXED if precision_name == "mixed" else hb.PRECISION_FP64
    dtype = torch.complex64 if precision_name == "mixed" else torch.complex128
    n = sp.global_dof_count
    ref = torch.empty((n, n), dtype=dtype, device=hb.device())
    hb.regular_assemble_symmetric("helmholtz_single_layer", dsp, rule, 2.5, precision, ref)
    pad = 5
    big = torch.full((n, n + pad), 7.0 + 3.0j, dtype=dtype, device=hb.device())
    hb.check(hb.lib().bem_regular_assemble_symmetric(
        hb.OPERATORS["helmholtz_single_layer"], ctypes.byref(dsp.struct), ctypes.byref(rule), 2.5, 0.0, precision,
        float(hb.DEFAULT_NEAR_FIELD_FACTOR), 0, hb.ptr(big), n + pad, hb._dtype_code(big), hb.stream_handle()))
    torch.cuda.synchronize()
    assert bool((big[:, :n] == ref).all())
    assert bool((big[:, n:] == 7.0 + 3.0j).all())
    assert bool((ref == ref.T).all())
    assert bool((torch.diagonal(ref) == 0).all())
