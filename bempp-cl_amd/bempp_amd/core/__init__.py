"""Core assemblers and the HIP device interface."""
