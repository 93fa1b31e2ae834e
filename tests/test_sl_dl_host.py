"""Host-side rules of the fused single/double layer assembly (no GPU)."""


def test_can_fuse_sl_dl():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.api.operators.boundary import helmholtz, laplace
    from bempp_amd.core.dense_assembler import can_fuse_sl_dl

    g = icosphere(1)
    p1 = function_space(g, "P", 1)
    dp0 = function_space(g, "DP", 0)

    def t(op):
        return (op.descriptor, op.domain, op.dual_to_range)

    sl = helmholtz.single_layer(p1, p1, p1, 2.0, precision="single")
    dl = helmholtz.double_layer(p1, p1, p1, 2.0, precision="single")
    assert can_fuse_sl_dl(t(sl), t(dl))
    assert not can_fuse_sl_dl(t(dl), t(sl))  # order: single layer first
    assert not can_fuse_sl_dl(t(sl), t(helmholtz.double_layer(p1, p1, p1, 3.0, precision="single")))
    assert not can_fuse_sl_dl(t(sl), t(helmholtz.double_layer(p1, p1, p1, 2.0, precision="double")))
    assert not can_fuse_sl_dl(t(sl), t(helmholtz.double_layer(dp0, dp0, dp0, 2.0, precision="single")))
    assert not can_fuse_sl_dl(t(sl), t(helmholtz.adjoint_double_layer(p1, p1, p1, 2.0, precision="single")))
    assert not can_fuse_sl_dl(t(laplace.single_layer(p1, p1, p1)), t(dl))
    assert can_fuse_sl_dl(t(laplace.single_layer(p1, p1, p1)), t(laplace.double_layer(p1, p1, p1)))


def test_fusion_key_includes_the_traversal_option():
    """A single and a double layer that differ only in parameters.hip.symmetric are not fused:
    the fused pass would run both with the single layer's traversal."""
    import bempp_amd.api as bempp
    from bempp_amd.api.assembly.boundary_operator import _fusion_key
    from bempp_amd.api.operators.boundary import helmholtz
    from bempp_amd.api.utils.parameters import DefaultParameters

    space = bempp.function_space(bempp.shapes.icosphere(1), "P", 1)
    p1, p2 = DefaultParameters(), DefaultParameters()
    p2.hip.symmetric = not p1.hip.symmetric
    sl = helmholtz.single_layer(space, space, space, 2.0, parameters=p1)
    dl1 = helmholtz.double_layer(space, space, space, 2.0, parameters=p1)
    dl2 = helmholtz.double_layer(space, space, space, 2.0, parameters=p2)
    assert _fusion_key(sl) == _fusion_key(dl1)
    assert _fusion_key(sl) != _fusion_key(dl2)


def test_release_drops_only_the_given_grid_and_spaces():
    import bempp_amd.api as bempp
    from bempp_amd.core import blocking, near_field, release

    g1, g2 = bempp.shapes.icosphere(1), bempp.shapes.icosphere(1)
    s1, s2 = bempp.function_space(g1, "P", 1), bempp.function_space(g2, "P", 1)
    for s in (s1, s2):
        blocking.make_blocks(s, 16, 20)
        near_field.pruned_candidates(s, s, 1.5, False)
    ids = lambda cache: {p for key in cache for p in (key if isinstance(key, tuple) else (key,)) if isinstance(p, str)}
    assert {s1.id, s2.id} <= ids(blocking._CACHE) & ids(near_field._CACHE)
    release(g1, (s1,))
    assert s1.id not in ids(blocking._CACHE) | ids(near_field._CACHE)
    assert s2.id in ids(blocking._CACHE) and s2.id in ids(near_field._CACHE)
    release(None, (s2,))
    assert s2.id not in ids(blocking._CACHE) | ids(near_field._CACHE)
