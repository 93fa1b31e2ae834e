"""The C-ABI library: it loads without a GPU and exports every entry point the header declares."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bempp_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["bem_regular_assemble", "bem_singular_assemble", "bem_potential_evaluate",
                     "bem_potential_workspace_size", "bem_fill_zero", "bem_last_error", "bem_version",
                     "bem_kbatch_assemble"]:
        assert required in names


def test_library_exports_all_symbols():
    from bempp_amd.core import hip_backend

    path = hip_backend.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.bem_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.bem_version()


def test_error_path_without_gpu():
    """Invalid arguments are rejected on the host before any device work."""
    from bempp_amd.core import hip_backend

    L = hip_backend.lib()
    rule = hip_backend.BemQuadRule()
    rule.n_points = 0
    sp = hip_backend.BemSpace()
    rc = L.bem_regular_assemble(0, ctypes.byref(sp), ctypes.byref(sp), 1, ctypes.byref(rule), 0.0, 0.0, 0, 2.0, 1,
                                ctypes.c_void_p(16), 1, 1, 0, 0, None)
    assert rc == 1
    assert b"quadrature" in L.bem_last_error()
    rc = L.bem_potential_evaluate(2, ctypes.byref(sp), ctypes.c_void_p(16), 1, ctypes.c_void_p(16),
                                  ctypes.byref(rule), 0.0, 0.0, 0, ctypes.c_void_p(16), 3, ctypes.c_void_p(16), 0, None)
    assert rc == 2  # adjoint double layer has no potential: KeyError class
    ks = (ctypes.c_double * 1)(1.0)
    rc = L.bem_kbatch_assemble(ctypes.byref(sp), ctypes.byref(sp), 1, ctypes.byref(rule), 1, ks, 1, 7, 2.0, None,
                               ctypes.c_void_p(16), 1, 1, 2, 0, None, 0, None, None, None, None, None)
    assert rc == 1
    assert b"null" in L.bem_last_error()
    cfg = hip_backend.BemKBatchConfig(16, 4)
    rc = L.bem_kbatch_assemble(ctypes.byref(sp), ctypes.byref(sp), 1, ctypes.byref(rule), 65, ks, 1, 7, 2.0,
                               ctypes.byref(cfg), ctypes.c_void_p(16), 1, 1, 2, 0, None, 0, None, None, None, None, None)
    assert rc != 0
    # symmetric pass: missing outputs / blocks and a malformed work list are rejected on the host
    tb = hip_backend.BemDofBlocks()
    rc = L.bem_regular_assemble_sl_dl_symmetric(3, ctypes.byref(sp), ctypes.byref(tb), ctypes.byref(tb), 1, None, 4,
                                                ctypes.byref(rule), 1.0, 0.0, 1, 2.0, 0, ctypes.c_void_p(16), None, 4,
                                                2, None)
    assert rc == 1
    assert b"null" in L.bem_last_error()
    rc = L.bem_regular_assemble_sl_dl_symmetric(3, ctypes.byref(sp), ctypes.byref(tb), ctypes.byref(tb), 0, None, 8,
                                                ctypes.byref(rule), 1.0, 0.0, 1, 2.0, 0, ctypes.c_void_p(16),
                                                ctypes.c_void_p(16), 4, 2, None)
    assert rc == 1  # ld < n_dofs
    rc = L.bem_dense_add_transpose(ctypes.c_void_p(16), 2, 8, 4, None)
    assert rc == 1


def test_noalloc_potential_refuses_oversize_without_gpu():
    """The allocation-free potential refuses a workspace beyond its static buffer on the host,
    before it touches a device (only n_elements and the rule's size enter the check)."""
    from bempp_amd.core import hip_backend

    L = hip_backend.lib()
    rule = hip_backend.BemQuadRule()
    rule.n_points = 6
    sp = hip_backend.BemSpace()
    sp.n_elements = 1 << 22
    n_points = 1 << 20
    need = L.bem_potential_workspace_size(3, ctypes.byref(sp), ctypes.byref(rule), n_points, 1)
    assert need > (64 << 20)
    dummy = ctypes.c_void_p(256)
    rc = L.bem_potential_evaluate_noalloc(3, ctypes.byref(sp), dummy, n_points, dummy, ctypes.byref(rule), 1.0,
                                          0.0, 1, dummy, 2, None)
    assert rc == 4
    assert b"scratch" in L.bem_last_error()


@pytest.mark.parametrize("n_elements,n_points", [(20, 0), (0, 100), (0, 0), (1, 1)])
def test_potential_workspace_size_empty_sets(n_elements, n_points):
    """The workspace plan of bem_potential_evaluate is computed on the host for empty point sets
    and empty spaces as well (a plan that divided by the point-block count once crashed here)."""
    from bempp_amd.core import hip_backend

    L = hip_backend.lib()
    rule = hip_backend.BemQuadRule()
    rule.n_points = 6
    sp = hip_backend.BemSpace()
    sp.n_elements = n_elements
    for op in (0, 3, 6):
        need = L.bem_potential_workspace_size(op, ctypes.byref(sp), ctypes.byref(rule), n_points, 1)
        assert need >= 256
