"""Candidate near-field element pairs for the mixed-precision mode.

In single precision the FP32 regular kernels skip the element pairs whose centroid
distance is below factor * (diameter_test + diameter_trial) / 2 and the FP64 near-field
kernel computes them. The device applies that test itself; the host only supplies a
superset of candidates: all (test, trial) support-element pairs with centroid distance
below factor * max diameter (times a safety margin), found with a k-d tree.
"""

import numpy as _np

_CACHE = {}


def near_candidates(test_space, trial_space, factor):
    """int32 array (n, 2) of candidate (test element, trial element) pairs."""
    from scipy.spatial import cKDTree

    key = (test_space.id, trial_space.id, float(factor))
    if key in _CACHE:
        return _CACHE[key]
    tg, sg = test_space.grid, trial_space.grid
    te = test_space.support_elements.astype(_np.int64)
    se = trial_space.support_elements.astype(_np.int64)
    dmax = max(float(tg.diameters[te].max()) if len(te) else 0.0, float(sg.diameters[se].max()) if len(se) else 0.0)
    radius = factor * dmax * 1.001 + 1e-12
    tree = cKDTree(sg.centroids[se])
    lists = tree.query_ball_point(tg.centroids[te], r=radius)
    counts = _np.fromiter((len(x) for x in lists), dtype=_np.int64, count=len(lists))
    rows = _np.repeat(te, counts)
    cols = se[_np.concatenate([_np.asarray(x, dtype=_np.int64) for x in lists])] if counts.sum() else _np.zeros(0, _np.int64)
    pairs = _np.ascontiguousarray(_np.stack([rows, cols], axis=1).astype(_np.int32))
    _CACHE[key] = pairs
    return pairs


def pruned_candidates(test_space, trial_space, factor, unordered=False):
    """The candidates without the pairs the device would reject for certain.

    Drops pairs that share a vertex (same grid: the singular kernel takes them) and pairs whose
    centroid distance exceeds factor * (diam_t + diam_f) / 2 by more than a 5e-5 relative margin
    (the device applies the exact test to what remains). unordered: keep each pair of one space
    with itself once, as (e, f) with e < f (the symmetric pass computes both orientations).
    """
    key = ("pruned", test_space.id, trial_space.id, float(factor), bool(unordered))
    if key in _CACHE:
        return _CACHE[key]
    pairs = near_candidates(test_space, trial_space, factor)
    e, f = pairs[:, 0].astype(_np.int64), pairs[:, 1].astype(_np.int64)
    tg, sg = test_space.grid, trial_space.grid
    keep = _np.ones(len(pairs), dtype=bool)
    if unordered:
        keep &= e < f
    if tg == sg:
        ve, vf = tg.elements[:, e], tg.elements[:, f]
        keep &= ~(ve[:, None, :] == vf[None, :, :]).any(axis=(0, 1))
    d = tg.centroids[e] - sg.centroids[f]
    lim = factor * 0.5 * (tg.diameters[e] + sg.diameters[f])
    keep &= (d * d).sum(axis=1) < lim * lim * (1.0 + 1e-4)
    out = _np.ascontiguousarray(pairs[keep])
    _CACHE[key] = out
    return out


_DEVICE_CACHE = {}


def device_near_candidates(test_space, trial_space, factor, unordered=False):
    import torch
    from bempp_amd.core import hip_backend as hb

    key = (test_space.id, trial_space.id, float(factor), bool(unordered))
    if key not in _DEVICE_CACHE:
        _DEVICE_CACHE[key] = torch.from_numpy(pruned_candidates(test_space, trial_space, factor, unordered)).to(
            hb.device())
    return _DEVICE_CACHE[key]


def clear_cache():
    _CACHE.clear()
    _DEVICE_CACHE.clear()
