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
