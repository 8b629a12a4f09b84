"""The C-ABI library: it loads, exports every entry point include/catana_hip.h
declares, and its host-only functions behave (no device work here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from catana_amd import capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(sfbhip_\w+)\s*\(", src))
    return names


def test_every_declared_symbol_is_exported():
    L = capi.lib()
    declared = _declared()
    assert "sfbhip_direct" in declared and "sfbhip_healpix" in declared
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared == set(capi.EXPORTED)


def test_default_params():
    p = capi.Params()
    capi.check(capi.lib().sfbhip_default_params(ctypes.byref(p)), "defaults")
    assert p.precision == capi.PREC_FP64
    assert p.bessel == capi.BESSEL_AUTO
    assert p.zeros == capi.ZEROS_REFERENCE
    assert p.alias_compat == 1
    assert p.mixed_fp64_below == 2
    assert p.mixed_operands == capi.OPERANDS_SPLIT_FP16
    assert p.mixed_chunks == 4
    assert p.mixed_kernel == capi.MIXED_KERNEL_AUTO
    # two ints, three doubles, nine ints, padded to the struct's 8-byte alignment
    assert ctypes.sizeof(capi.Params) == 72


def test_flmn_count():
    for lmax, nmax in [(0, 5), (1, 1), (16, 15), (256, 160)]:
        assert capi.flmn_count(lmax, nmax) == sum((l + 1) * nmax for l in range(lmax))


@pytest.mark.parametrize("nside", [1, 2, 16, 256, 8192])
def test_ang2pix_ring_matches_oracle(nside):
    rng = np.random.default_rng(nside)
    th = np.arccos(rng.uniform(-1, 1, 20000))
    ph = rng.uniform(0, 2 * np.pi, 20000)
    # poles, the sin(theta) branch boundary, cap/belt boundaries and phi wrap
    th = np.concatenate([th, [0.0, 1e-9, 0.0099, 0.01, 0.0101, np.pi - 0.0101, np.pi - 1e-9, np.pi,
                              np.arccos(2 / 3), np.arccos(-2 / 3), np.pi / 2, np.pi / 2]])
    ph = np.concatenate([ph, [0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 0.3, 1.1, 2.2, 2 * np.pi, -1e-12]])
    L = capi.lib()
    out = ctypes.c_int64()
    got = np.empty(len(th), np.int64)
    for i in range(len(th)):
        capi.check(L.sfbhip_ang2pix_ring(nside, float(th[i]), float(ph[i]), ctypes.byref(out)), "ang2pix")
        got[i] = out.value
    np.testing.assert_array_equal(got, O.ang2pix_ring(nside, th, ph))
    assert got.min() >= 0 and got.max() < 12 * nside * nside


def test_ang2pix_ring_rejects_bad_nside():
    out = ctypes.c_int64()
    for ns in (0, 3, 255, 16384):
        with pytest.raises(ValueError, match="power of 2"):
            capi.check(capi.lib().sfbhip_ang2pix_ring(ns, 0.5, 0.5, ctypes.byref(out)), "ang2pix")


def test_bessel_zeros_host():
    # SphericalBesselZeros known answers (SphericalBesselZeros_test.cpp:9-15,
    # test/python/besseltools_test.py:11-19) and the oracle's zeros
    z0 = capi.bessel_zeros(0, 3)
    np.testing.assert_allclose(z0, [np.pi, 2 * np.pi, 3 * np.pi], rtol=1e-15)
    assert abs(capi.bessel_zeros(5, 10)[9] - 38.8836) < 1e-4
    for l in (0, 1, 7, 64, 255):
        np.testing.assert_allclose(capi.bessel_zeros(l, 40), O.bessel_zeros(l, 40), rtol=1e-14)
    with pytest.raises(ValueError):
        capi.bessel_zeros(-1, 3)
