"""Point containers (catana.basictypes: python/py_basictypes.cpp:22-154)."""
from ._catana import PixelizedPointContainer, Point, PointContainer  # noqa: F401

__all__ = ["Point", "PointContainer", "PixelizedPointContainer"]
