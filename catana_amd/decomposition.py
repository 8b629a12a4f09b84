"""SFB decomposition (catana.decomposition: python/py_decomposition.cpp:27-153)."""
from ._catana import (  # noqa: F401
    KClkk, _sfb_raw, _sfb_reverse, _sfb_reverse_fft, compute_kln, sfb_decomposition, sfb_decomposition_healpix,
    sfb_decomposition_reverse, sfb_decomposition_reverse_batched)

__all__ = ["KClkk", "sfb_decomposition", "sfb_decomposition_reverse", "sfb_decomposition_reverse_batched",
           "sfb_decomposition_healpix", "compute_kln"]
