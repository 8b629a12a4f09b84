"""catana_amd -- MI355X-native spherical Fourier-Bessel decomposition.

A drop-in for the SFB hot path of CatAna (catana.sfb_decomposition and its
HEALPix variant).  All arithmetic runs in libcatana_hip.so (hand-written HIP
for gfx950) behind the C-ABI in include/catana_hip.h.
"""
__version__ = "0.1.0"
