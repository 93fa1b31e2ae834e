"""Singular pair lists and Sauter-Schwab point tables for the device singular kernel.

Counterpart of bempp_cl/core/singular_assembler.py:69-448. The host builds, once per
(grid, order, supports):
* the concatenated reference points of all rules and their remaps (coincident rule, the
  6 shared-edge remaps in the order (0,1), (1,0), (1,2), (2,1), (0,2), (2,0), the 3 shared-
  vertex remaps), test and trial, and the concatenated weights (singular_assembler.py:278-320,
  372-404);
* per pair: test element, trial element, test point offset, trial point offset, weight
  offset and point count, with edge offsets from the table [[-1, 0, 4], [1, -1, 2], [5, 3, -1]]
  indexed by the local vertex pair (singular_assembler.py:322-337, 406-448).
Pairs are coincident pairs, then edge-adjacent, then vertex-adjacent, filtered by the
supports (singular_assembler.py:166-186). The arrays are uploaded to HBM and reused.
"""

import numpy as _np

from bempp_amd.api.integration import duffy_galerkin as _duffy

_EDGE_REMAPS = [(0, 1), (1, 0), (1, 2), (2, 1), (0, 2), (2, 0)]
_EDGE_OFFSET_INDEX = _np.array([[-1, 0, 4], [1, -1, 2], [5, 3, -1]])

_CACHE = {}


class SingularRule(object):
    """Host arrays of all singular pairs of a grid for given supports."""

    def __init__(self, grid, order, test_support, trial_support):
        co_t, co_s, co_w = _duffy.rule(order, "coincident")
        ed_t, ed_s, ed_w = _duffy.rule(order, "edge_adjacent")
        ve_t, ve_s, ve_w = _duffy.rule(order, "vertex_adjacent")
        nco, ned, nve = len(co_w), len(ed_w), len(ve_w)

        self.test_points = _np.hstack(
            [co_t]
            + [_duffy.remap_points_shared_edge(ed_t, a, b) for a, b in _EDGE_REMAPS]
            + [_duffy.remap_points_shared_vertex(ve_t, v) for v in range(3)]
        )
        self.trial_points = _np.hstack(
            [co_s]
            + [_duffy.remap_points_shared_edge(ed_s, a, b) for a, b in _EDGE_REMAPS]
            + [_duffy.remap_points_shared_vertex(ve_s, v) for v in range(3)]
        )
        self.weights = _np.hstack([co_w, ed_w, ve_w])

        coincident = _np.flatnonzero(test_support & trial_support)
        ea = grid.edge_adjacency
        ea = ea[:, test_support[ea[0]] & trial_support[ea[1]]]
        va = grid.vertex_adjacency
        va = va[:, test_support[va[0]] & trial_support[va[1]]]

        edge_offsets = nco + ned * _EDGE_OFFSET_INDEX
        vertex_offsets = nco + 6 * ned + nve * _np.arange(3)

        n = len(coincident) + ea.shape[1] + va.shape[1]
        pairs = _np.zeros((n, 6), dtype=_np.int32)
        i0, i1 = len(coincident), len(coincident) + ea.shape[1]
        pairs[:i0, 0] = coincident
        pairs[:i0, 1] = coincident
        pairs[:i0, 2] = 0
        pairs[:i0, 3] = 0
        pairs[:i0, 4] = 0
        pairs[:i0, 5] = nco
        pairs[i0:i1, 0] = ea[0]
        pairs[i0:i1, 1] = ea[1]
        pairs[i0:i1, 2] = edge_offsets[ea[2], ea[3]]
        pairs[i0:i1, 3] = edge_offsets[ea[4], ea[5]]
        pairs[i0:i1, 4] = nco
        pairs[i0:i1, 5] = ned
        pairs[i1:, 0] = va[0]
        pairs[i1:, 1] = va[1]
        pairs[i1:, 2] = vertex_offsets[va[2]]
        pairs[i1:, 3] = vertex_offsets[va[3]]
        pairs[i1:, 4] = nco + ned
        pairs[i1:, 5] = nve
        self.pairs = pairs
        self.counts = {"coincident": i0, "edge_adjacent": i1 - i0, "vertex_adjacent": n - i1}

    @property
    def number_of_pairs(self):
        return self.pairs.shape[0]


class DeviceSingularRule(object):
    """SingularRule arrays in HBM."""

    def __init__(self, rule):
        import torch
        from bempp_amd.core import hip_backend

        dev = hip_backend.device()
        self.host = rule
        self.pairs = torch.from_numpy(rule.pairs).to(dev)
        self.test_points = torch.from_numpy(_np.ascontiguousarray(rule.test_points.T)).to(dev)
        self.trial_points = torch.from_numpy(_np.ascontiguousarray(rule.trial_points.T)).to(dev)
        self.weights = torch.from_numpy(_np.ascontiguousarray(rule.weights)).to(dev)


def singular_rule(grid, order, test_support, trial_support, on_device=True):
    # the pairs and the reference-coordinate points depend on the topology only
    key = (grid.topology_id, int(order), test_support.tobytes(), trial_support.tobytes(), on_device)
    if key not in _CACHE:
        rule = SingularRule(grid, order, _np.asarray(test_support, bool), _np.asarray(trial_support, bool))
        _CACHE[key] = DeviceSingularRule(rule) if on_device else rule
    return _CACHE[key]


def clear_cache():
    _CACHE.clear()
