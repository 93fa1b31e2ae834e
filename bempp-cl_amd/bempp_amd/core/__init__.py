"""Core assemblers and the HIP device interface."""


def clear_caches():
    """Drop every host and device cache (grids, spaces, blocks, near-field and singular pairs)."""
    from bempp_amd.core import blocking, hip_backend, near_field, singular_assembler

    hip_backend.clear_caches()
    blocking.clear_cache()
    near_field.clear_cache()
    singular_assembler.clear_cache()


def _drop_keys(cache, ids):
    for key in list(cache):
        parts = key if isinstance(key, tuple) else (key,)
        if any(part in ids for part in parts if isinstance(part, str)):
            del cache[key]


def release(grid=None, spaces=()):
    """Drop the cached host and device data of one grid and of some spaces.

    Device copies of geometry and dof maps, dof blocks and near-field candidates are cached by
    grid and space id, so that repeated assemblies on one mesh reuse them. A caller that moves
    the vertices (a new grid per step, as in shape optimisation) releases the previous grid and
    its spaces, so the caches do not grow by one mesh per step. Singular pair tables depend only
    on the topology and are kept. Released objects stay usable: their data is rebuilt on demand.
    """
    from bempp_amd.core import blocking, hip_backend, near_field

    ids = {s.id for s in spaces if s is not None}
    if grid is not None:
        ids.add(grid.id)
    if not ids:
        return
    for cache in (hip_backend._GRID_CACHE, hip_backend._SPACE_CACHE, blocking._CACHE, blocking._DEVICE_CACHE,
                  near_field._CACHE, near_field._DEVICE_CACHE):
        _drop_keys(cache, ids)
