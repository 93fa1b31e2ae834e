"""Host-side blocking for the LDS-tiled scatter and the near-field candidate lists.

The tiled kernel adds, per (test block, trial block), sum over the block's element pairs of
m_i m_j L_ef[i, j] into tile[slot(e, i), slot(f, j)] and then the tile into
A[dofs_t[a], dofs_s[b]]. Emulating that in numpy with random local blocks must give the
same matrix as the reference's direct scatter (np.add.at over local2global), which pins
the slot/dof bookkeeping independently of the GPU.
"""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ico3():
    from bempp_amd.api.grid.shapes import icosphere

    return icosphere(3)


@pytest.mark.parametrize("kind", [("P", 1), ("RWG", 0), ("DP", 0)])
@pytest.mark.parametrize("limits", [(16, 20), (256, 160)])
def test_blocks_are_consistent(ico3, kind, limits):
    from bempp_amd.api import function_space
    from bempp_amd.core.blocking import make_blocks

    sp = function_space(ico3, *kind)
    me, md = limits
    b = make_blocks(sp, me, md)
    assert sorted(b.order.tolist()) == sorted(sp.support_elements.tolist())
    assert b.start[0] == 0 and b.start[-1] == len(b.order)
    assert np.all(np.diff(b.start) >= 1) and np.all(np.diff(b.start) <= me)
    assert np.all(b.ndofs <= md)
    nshape = sp.number_of_shape_functions
    for blk in range(b.n_blocks):
        for pos in range(b.start[blk], b.start[blk + 1]):
            e = b.order[pos]
            for i in range(nshape):
                if sp.local_multipliers[e, i] != 0:
                    assert b.dofs[blk, b.slots[pos, i]] == sp.local2global[e, i]
                    assert b.slots[pos, i] < b.ndofs[blk]


def test_tiled_scatter_equals_direct_scatter():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.core.blocking import make_blocks

    grid = icosphere(2)
    sp = function_space(grid, "P", 1)
    F = grid.number_of_elements
    rng = np.random.default_rng(3)
    L = rng.standard_normal((F, F, 3, 3))
    m = sp.local_multipliers
    l2g = sp.local2global
    N = sp.global_dof_count
    direct = np.zeros((N, N))
    rows = np.repeat(l2g[:, None, :, None], 3, axis=3) * np.ones((1, F, 1, 1), dtype=np.int64)
    cols = np.repeat(l2g[None, :, None, :], 3, axis=2) * np.ones((F, 1, 1, 1), dtype=np.int64)
    vals = L * m[:, None, :, None] * m[None, :, None, :]
    np.add.at(direct, (rows.ravel(), cols.ravel()), vals.ravel())

    tb = make_blocks(sp, 16, 20)
    sb = make_blocks(sp, 256, 160)
    inc = sb.build_incidence(m, 12)
    # incidence lists name exactly the (element, shape function) pairs of each slot
    for bs in range(sb.n_blocks):
        for s in range(sb.ndofs[bs]):
            codes = [c for c in inc[bs, s] if c >= 0]
            assert codes
            for c in codes:
                pos = sb.start[bs] + c // 3
                assert sb.slots[pos, c % 3] == s and m[sb.order[pos], c % 3] != 0
    tiled = np.zeros((N, N))
    for bt in range(tb.n_blocks):
        for bs in range(sb.n_blocks):
            tile = np.zeros((20, 160))
            for pt in range(tb.start[bt], tb.start[bt + 1]):
                e = tb.order[pt]
                for ps in range(sb.start[bs], sb.start[bs + 1]):
                    f = sb.order[ps]
                    for i in range(3):
                        for j in range(3):
                            if m[e, i] * m[f, j] != 0:
                                tile[tb.slots[pt, i], sb.slots[ps, j]] += L[e, f, i, j] * m[e, i] * m[f, j]
            nt, ns = tb.ndofs[bt], sb.ndofs[bs]
            np.add.at(tiled, (np.repeat(tb.dofs[bt, :nt], ns), np.tile(sb.dofs[bs, :ns], nt)), tile[:nt, :ns].ravel())
    np.testing.assert_allclose(tiled, direct, rtol=1e-12, atol=1e-12)


def test_near_candidates_superset(ico3):
    from bempp_amd.api import function_space
    from bempp_amd.core.near_field import near_candidates

    sp = function_space(ico3, "DP", 0)
    factor = 2.0
    cand = near_candidates(sp, sp, factor)
    cand_set = set(map(tuple, cand.tolist()))
    c = ico3.centroids
    d = ico3.diameters
    dist2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    lim = factor * 0.5 * (d[:, None] + d[None, :])
    need = np.argwhere(dist2 < lim**2 * (1 + 1e-9))
    missing = [tuple(p) for p in need.tolist() if tuple(p) not in cand_set]
    assert not missing
