"""ofri: host side of the MI355X OpticalFlow-RI implementation.

The drop-in modules next to this package (HornSchunck, PhysicsBasedOpticalFlowLiuShen,
GenericPyramidalOpticalFlow, gaussian_filter, GaussianKernelBitExact, ...) keep
the reference's Python API and call libofri.so through this package.
"""
from ._capi import OfriError, load  # noqa: F401
from .device import Context, DeviceArray, context  # noqa: F401

__all__ = ["OfriError", "load", "Context", "DeviceArray", "context"]
