"""catana_amd -- MI355X-native spherical Fourier-Bessel decomposition.

A drop-in for the SFB hot path of CatAna (catana.sfb_decomposition and its
HEALPix variant).  All arithmetic runs in libcatana_hip.so (hand-written HIP
for gfx950) behind the C-ABI in include/catana_hip.h; the Python names come
from the pybind11 module _catana, the same names as catana's basictypes and
decomposition modules (python/catana/__init__.py:1-4 re-exports them the same
way), plus catana_amd.besseltools (the double spherical-Bessel integrals of the
shot-noise step), catana_amd.io (sources, sinks and point filters, filtering on the
device), Analyzer and shot_noise().  The names are resolved on first use, so that ``import catana_amd``
works before the extension is built; using one without the built extension
raises ImportError.
"""
__version__ = "0.1.0"

_API = (
    "Point", "PointContainer", "PixelizedPointContainer", "KClkk", "sfb_decomposition",
    "sfb_decomposition_reverse", "sfb_decomposition_reverse_batched", "sfb_decomposition_healpix",
    "compute_kln", "_sfb_raw", "_sfb_reverse", "_sfb_reverse_fft", "FunctionInterpolator", "shot_noise",
    "Analyzer",
)
__all__ = [n for n in _API if not n.startswith("_")]


def __getattr__(name):
    if name in _API:
        from . import _catana
        return getattr(_catana, name)
    if name in ("besseltools", "io"):
        import importlib
        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(f"module 'catana_amd' has no attribute {name!r}")


def init_random(seed=None):
    """Seed of the random point filters and subsets (python/catana/__init__.py:7-21):
    every draw is a hash of (seed, filter, running record count)."""
    from . import _catana
    if seed is None:
        _catana.init_random()
    else:
        _catana.init_random(int(seed))
