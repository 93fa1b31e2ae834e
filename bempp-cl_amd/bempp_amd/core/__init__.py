"""Core assemblers and the HIP device interface."""


def clear_caches():
    """Drop every host and device cache (grids, spaces, blocks, near-field and singular pairs)."""
    from bempp_amd.core import blocking, hip_backend, near_field, singular_assembler

    hip_backend.clear_caches()
    blocking.clear_cache()
    near_field.clear_cache()
    singular_assembler.clear_cache()
