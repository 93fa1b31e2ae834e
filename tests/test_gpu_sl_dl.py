"""Single + double layer assembled in one pass (bem_regular/near/singular_assemble_sl_dl).

Each fused matrix must equal the separately assembled operator: FP64 fused vs FP64 separate
within 1e-13 (and vs the FP64 oracle within 1e-12), mixed fused vs FP64 within 1e-6.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    import torch

    return (torch.linalg.norm(a.to(b.dtype) - b) / torch.linalg.norm(b)).item()


@pytest.fixture(scope="module")
def ico2():
    from bempp_amd.api.grid.shapes import icosphere

    return icosphere(2)


CASES = [
    ("helmholtz", 4.0, ("P", 1), ("P", 1)),
    ("helmholtz", 3.0 + 0.5j, ("P", 1), ("P", 1)),
    ("helmholtz", 2.0, ("DP", 0), ("DP", 0)),
    ("helmholtz", 5.0, ("DP", 0), ("P", 1)),
    ("helmholtz", 2.5, ("P", 1), ("DP", 1)),
    ("laplace", None, ("P", 1), ("P", 1)),
    ("laplace", None, ("DP", 0), ("DP", 0)),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}_{c[1]}_{c[2][0]}{c[2][1]}_{c[3][0]}{c[3][1]}" for c in CASES])
def test_fused_equals_separate(case, ico2, default_parameters, hip_available):
    import bempp_amd.api.operators.boundary as B
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import hip_dense_assemble, hip_dense_assemble_sl_dl

    fam, k, dk, tk = case
    dom = function_space(ico2, *dk)
    dual = function_space(ico2, *tk)
    mod = getattr(B, fam)
    args = (dom, dom, dual) if k is None else (dom, dom, dual, k)
    for precision in ("double", "single"):
        sl = mod.single_layer(*args, parameters=default_parameters, precision=precision).descriptor
        dl = mod.double_layer(*args, parameters=default_parameters, precision=precision).descriptor
        a1, a2 = hip_dense_assemble_sl_dl(dom, dual, default_parameters, sl, dl)
        sl64 = mod.single_layer(*args, parameters=default_parameters, precision="double").descriptor
        dl64 = mod.double_layer(*args, parameters=default_parameters, precision="double").descriptor
        r1 = hip_dense_assemble(dom, dual, default_parameters, sl64)
        r2 = hip_dense_assemble(dom, dual, default_parameters, dl64)
        tol = 1e-13 if precision == "double" else 1e-6
        assert rel(a1, r1) < tol, precision
        assert rel(a2, r2) < tol, precision


def test_fused_fp64_vs_oracle(ico2, default_parameters, hip_available):
    from bempp_amd.api import function_space
    from bempp_amd.api.operators.boundary import helmholtz
    from bempp_amd.core.dense_assembler import hip_dense_assemble_sl_dl
    from oracle import oracle as O

    p1 = function_space(ico2, "P", 1)
    k = 4.0
    sl = helmholtz.single_layer(p1, p1, p1, k, parameters=default_parameters, precision="double").descriptor
    dl = helmholtz.double_layer(p1, p1, p1, k, parameters=default_parameters, precision="double").descriptor
    a1, a2 = hip_dense_assemble_sl_dl(p1, p1, default_parameters, sl, dl)
    for name, a in (("helmholtz_single_layer", a1), ("helmholtz_double_layer", a2)):
        ref = O.assemble(name, ico2.vertices, ico2.elements, p1, p1, k)
        assert np.linalg.norm(a.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-12


def test_assemble_weak_forms_api(ico2, default_parameters, hip_available):
    from bempp_amd.api import function_space
    from bempp_amd.api.assembly.boundary_operator import assemble_weak_forms
    from bempp_amd.api.operators.boundary import helmholtz, laplace

    p1 = function_space(ico2, "P", 1)
    dp0 = function_space(ico2, "DP", 0)
    ops = [
        helmholtz.double_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="single"),
        laplace.single_layer(dp0, dp0, dp0, parameters=default_parameters),
        helmholtz.single_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="single"),
    ]
    forms = assemble_weak_forms(ops)
    assert ops[0]._cached is forms[0] and ops[2]._cached is forms[2]
    refs = [
        helmholtz.double_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="double").weak_form(),
        laplace.single_layer(dp0, dp0, dp0, parameters=default_parameters, precision="double").weak_form(),
        helmholtz.single_layer(p1, p1, p1, 4.0, parameters=default_parameters, precision="double").weak_form(),
    ]
    for f, r in zip(forms, refs):
        assert np.linalg.norm(f.to_dense() - r.to_dense()) / np.linalg.norm(r.to_dense()) < 1e-6
