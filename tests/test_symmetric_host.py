"""Host side of the symmetric single + double layer pass (no GPU): the nested blocking and
its work list visit every unordered element pair exactly once."""

import numpy as np
import pytest


# (level, tiles per item, space); RWG with the mixed electric field pass's 416-dof trial blocks
CASES = [(1, None, "P1"), (2, None, "P1"), (2, 3, "P1"), (3, None, "P1"), (3, 7, "P1"), (2, None, "RWG"),
         (3, None, "RWG"), (3, 5, "RWG")]


@pytest.mark.parametrize("level,run,kind", CASES)
def test_work_list_covers_each_unordered_pair_once(level, run, kind):
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.core.blocking import (RWG_MIXED_TRIAL_BLOCK_DOFS, SYM_RUN_ELEMENTS, SYM_TEST_TILE_DOFS,
                                         SYM_TEST_TILE_ELEMENTS, TRIAL_BLOCK_DOFS, make_symmetric_blocks)

    grid = icosphere(level)
    if kind == "RWG":
        p1 = function_space(grid, "RWG", 0)
        trial, tiles, owner, work = make_symmetric_blocks(p1, tiles_per_item=run,
                                                          trial_dofs=RWG_MIXED_TRIAL_BLOCK_DOFS, min_tiles_per_item=2)
        assert trial.ndofs.max() <= RWG_MIXED_TRIAL_BLOCK_DOFS
    else:
        p1 = function_space(grid, "P", 1)
        trial, tiles, owner, work = make_symmetric_blocks(p1, tiles_per_item=run)
        assert trial.ndofs.max() <= TRIAL_BLOCK_DOFS
    # a work item's run of test tiles fits the kernel's LDS copy of the run
    assert (tiles.start[work[:, 2]] - tiles.start[work[:, 1]]).max() <= SYM_RUN_ELEMENTS
    n = len(trial.order)
    assert np.array_equal(tiles.order, trial.order)
    assert np.diff(tiles.start).max() <= SYM_TEST_TILE_ELEMENTS and tiles.ndofs.max() <= SYM_TEST_TILE_DOFS
    # every tile lies inside the block it is assigned to
    for t, b in enumerate(owner):
        assert trial.start[b] <= tiles.start[t] and tiles.start[t + 1] <= trial.start[b + 1]
    # the kernel's rule: in the tile's own block only trial positions after the test position
    count = np.zeros((n, n), dtype=np.int32)
    assert work.shape[1] == 3 and (work[:, 1] < work[:, 2]).all()
    for b, ta, tb in work:
        for t in range(ta, tb):
            t0, t1 = tiles.start[t], tiles.start[t + 1]
            s0, s1 = trial.start[b], trial.start[b + 1]
            pe = np.arange(t0, t1)[:, None]
            pf = np.arange(s0, s1)[None, :]
            take = np.ones((t1 - t0, s1 - s0), dtype=bool)
            if s0 <= t0 < s1:
                take = pf > pe
            np.add.at(count, (np.broadcast_to(pe, take.shape)[take], np.broadcast_to(pf, take.shape)[take]), 1)
    both = count + count.T
    off = ~np.eye(n, dtype=bool)
    assert (both[off] == 1).all()
    assert (np.diag(both) == 0).all()
    # slots of the tiles name the dofs of their elements
    l2g = p1.local2global
    for t in range(tiles.n_blocks):
        for pos in range(tiles.start[t], tiles.start[t + 1]):
            for i in range(3):
                assert tiles.dofs[t, tiles.slots[pos, i]] == l2g[trial.order[pos], i]


def test_symmetric_efie_predicate():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.core.dense_assembler import use_symmetric_efie

    g = icosphere(1)
    rwg, snc = function_space(g, "RWG", 0), function_space(g, "SNC", 0)
    assert use_symmetric_efie(rwg, snc)
    assert use_symmetric_efie(rwg, rwg)
    assert not use_symmetric_efie(rwg, function_space(icosphere(2), "SNC", 0))
    assert not use_symmetric_efie(function_space(g, "P", 1), function_space(g, "P", 1))
    part = function_space(g, "SNC", 0, support_elements=np.arange(40))
    assert not use_symmetric_efie(rwg, part)


def test_symmetric_pass_predicate():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere
    from bempp_amd.core.dense_assembler import use_symmetric_pass

    g = icosphere(1)
    p1 = function_space(g, "P", 1)
    assert use_symmetric_pass(p1, function_space(g, "P", 1))
    assert not use_symmetric_pass(p1, function_space(g, "DP", 1))
    assert not use_symmetric_pass(function_space(g, "DP", 0), function_space(g, "DP", 0))
    assert not use_symmetric_pass(p1, function_space(icosphere(2), "P", 1))


def test_upper_triangle_predicate():
    from bempp_amd.api import function_space
    from bempp_amd.core.dense_assembler import use_upper_triangle

    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(1)
    dp0, p1 = function_space(g, "DP", 0), function_space(g, "P", 1)
    assert use_upper_triangle(dp0, dp0, "helmholtz_single_layer")
    assert not use_upper_triangle(dp0, dp0, "helmholtz_double_layer")
    assert not use_upper_triangle(p1, p1, "laplace_single_layer")
    assert not use_upper_triangle(dp0, function_space(g, "DP", 0, support_elements=np.arange(10)), "laplace_single_layer")


def test_blockings_of_a_space_without_support_elements():
    """A space restricted to no elements (here RWG on one triangle: no interior edge) gives one
    empty block, one empty tile and one work item over it, which the kernels skip."""
    import bempp_amd.api as bempp
    from bempp_amd.core.blocking import RWG_MIXED_TRIAL_BLOCK_DOFS, make_blocks, make_symmetric_blocks

    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]).T
    rwg = bempp.function_space(bempp.Grid(v, np.array([[0, 1, 2]]).T), "RWG", 0)
    assert len(rwg.support_elements) == 0
    b = make_blocks(rwg, 256, RWG_MIXED_TRIAL_BLOCK_DOFS)
    assert b.n_blocks == 1 and b.ndofs.sum() == 0
    trial, tiles, owner, work = make_symmetric_blocks(rwg, trial_dofs=RWG_MIXED_TRIAL_BLOCK_DOFS,
                                                      min_tiles_per_item=2)
    assert len(trial.order) == 0 and tiles.ndofs.sum() == 0
    assert (tiles.start[work[:, 2]] - tiles.start[work[:, 1]]).max() == 0
