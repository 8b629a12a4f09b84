"""Spherical Bessel zeros and double spherical-Bessel integrals
(catana.besseltools: python/py_besseltools.cpp:12-80).

double_sbessel_integrator(w, l, r_max, k1, k2) computes
int_0^r_max w(r) j_l(k1 r) j_l(k2 r) dr on the device, vectorised over k1, k2;
w is a callable (evaluated at the quadrature nodes, in one call when it accepts
an array) or a catana_amd.FunctionInterpolator (interpolated on the device).
"""
from ._catana import besseltools as _bt

SphericalBesselZeros = _bt.SphericalBesselZeros
double_sbessel_integrator = _bt.double_sbessel_integrator
double_sbessel_integrals = _bt.double_sbessel_integrals

__all__ = ["SphericalBesselZeros", "double_sbessel_integrator", "double_sbessel_integrals"]
