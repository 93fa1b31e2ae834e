"""Dispatch to device interfaces (bempp_cl/core/dispatcher.py:4-55).

The reference accepts "numba" and "opencl"; this library provides the "hip" interface
(libbempp_hip.so on MI355X). Any other interface raises ValueError, as in the reference.
"""


def check_device_interface(device_interface):
    interface_type = str(device_interface).split("_")[0]
    if interface_type != "hip":
        raise ValueError("Device interface must be 'hip'.")


def dense_assembler_dispatcher(device_interface, *args, **kwargs):
    """Regular + singular dense assembly on the device; returns a device tensor."""
    check_device_interface(device_interface)
    from bempp_amd.core.dense_assembler import hip_dense_assemble

    return hip_dense_assemble(*args, **kwargs)


def potential_dispatcher(device_interface, *args, **kwargs):
    """Return an evaluator closure for a dense potential."""
    check_device_interface(device_interface)
    from bempp_amd.core.dense_potential_assembler import hip_potential_assembler

    return hip_potential_assembler(*args, **kwargs)
