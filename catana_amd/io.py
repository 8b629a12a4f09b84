"""Sources, sinks and point filters (catana.io: python/py_io.cpp:21-400).

The filters run on the device (sfbhip_filter_points): a FilterStream reads its
source in chunks, filters each chunk on the GPU and keeps the survivors in
device memory; the Analyzer decomposes them there without a host round trip.
"""
from ._catana import io as _io

_NAMES = [
    "init_random", "Source", "Sink", "Filter", "PointContainerSource", "GadgetSource", "CartesianTextSource",
    "SphericalTextSource", "SphericalTextSource_3dex", "RawBinarySource", "PointContainerSink",
    "PixelizedPointContainerSink", "CartesianTextSink", "SphericalTextSink", "SphericalTextSink_3dex",
    "RawBinarySink", "TophatRadialWindowFunctionFilter", "GaussianRadialWindowFunctionFilter",
    "GenericRadialWindowFunctionFilter", "AngularMaskFilter", "FilterStream",
]
globals().update({n: getattr(_io, n) for n in _NAMES})
__all__ = list(_NAMES)
