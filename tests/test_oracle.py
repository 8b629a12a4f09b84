"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

Each check is chosen so that a plausible mistake fails it: golden vectors from
independent implementations (mpmath, scipy, numpy.fft; tests/golden/make_golden.py),
the reference's own fixtures and known-answer tests, and invariants fixed by
the mathematics (addition theorem, rotation invariance).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from catana_amd import workloads

CF = 5535184878.03 / (2500 ** 3 * 4 / 3 * np.pi)  # test/python/decomposition_test.py:15


@pytest.fixture(scope="module")
def sf(golden_dir):
    return np.load(os.path.join(golden_dir, "special_functions.npz"))


# ---------------------------------------------------------------- j_l
def test_jl_golden(sf):
    a = np.array([O.lib().orc_sbessel_jl(int(l), float(x)) for l, x in zip(sf["jl_l"], sf["jl_x"])])
    v = sf["jl_v"]
    # scale: the local amplitude of j_l is ~ 1/x above the turning point; below it
    # the value itself sets the scale (relative accuracy)
    scale = np.where(sf["jl_x"] > sf["jl_l"], np.maximum(np.abs(v), 1.0 / sf["jl_x"]), np.abs(v))
    tiny = np.abs(v) < 1e-290
    err = np.abs(a - v)[~tiny] / scale[~tiny]
    assert err.max() < 1e-12, err.max()
    assert np.all(np.abs(a[tiny]) < 1e-280)


def test_jl_closed_forms():
    x = np.linspace(0.01, 300, 2000)
    np.testing.assert_allclose(O.sbessel_jl(0, x), np.sin(x) / x, rtol=1e-14, atol=1e-17)
    j1 = np.sin(x) / x ** 2 - np.cos(x) / x
    assert np.all(np.abs(O.sbessel_jl(1, x) - j1) <= 1e-13 * np.maximum(np.abs(j1), 1.0 / x))
    assert O.lib().orc_sbessel_jl(0, 0.0) == 1.0 and O.lib().orc_sbessel_jl(3, 0.0) == 0.0


# ---------------------------------------------------------------- P_l^m
def test_plm_golden(sf):
    a = np.array([O.lib().orc_sph_plm(int(l), int(m), float(x))
                  for l, m, x in zip(sf["pl_l"], sf["pl_m"], sf["pl_x"])])
    v = sf["pl_v"]
    scale = np.sqrt((2 * sf["pl_l"] + 1) / (4 * np.pi))
    assert np.max(np.abs(a - v) / scale) < 1e-12


def test_plm_addition_theorem():
    # sum_m |Y_lm|^2 = (2l+1)/(4 pi): |P_l^0|^2 + 2 sum_{m>=1} |P_l^m|^2
    for l in [0, 3, 30, 120, 255]:
        for x in [-0.97, -0.2, 0.0, 0.41, 0.9999]:
            p = np.array([O.lib().orc_sph_plm(l, m, x) for m in range(l + 1)])
            s = p[0] ** 2 + 2 * np.sum(p[1:] ** 2)
            assert abs(s / ((2 * l + 1) / (4 * np.pi)) - 1) < 1e-11


# ---------------------------------------------------------------- zeros / k_ln
def test_zeros_converged_golden(sf):
    for i, l in enumerate(sf["zeros_l"]):
        z = O.bessel_zeros(int(l), 20, O.ZEROS_CONVERGED)
        np.testing.assert_allclose(z, sf["zeros_true"][i], rtol=3e-15)


def test_zeros_reference_newton_golden(sf):
    """The reference's Newton loop stops after the first upward step
    (SphericalBesselZeros.cpp:59), so its zeros are not converged; the golden
    values come from the same loop on mpmath Bessel values."""
    worst_vs_true = 0.0
    for i, l in enumerate(sf["zeros_l"]):
        z = O.bessel_zeros(int(l), 20, O.ZEROS_REFERENCE)
        np.testing.assert_allclose(z, sf["zeros_reference"][i], rtol=1e-13)
        worst_vs_true = max(worst_vs_true, np.max(np.abs(z / sf["zeros_true"][i] - 1)))
    assert 1e-8 < worst_vs_true < 1e-4


def test_zeros_reference_known_answers():
    # SphericalBesselZeros_test.cpp:7-15 (Mathematica values)
    z0 = O.bessel_zeros(0, 3)
    np.testing.assert_allclose(z0, [3.141592653589791, 6.283185307179586, 9.42477796076938], rtol=1e-14)
    assert abs(O.bessel_zeros(5, 10)[9] / 38.8836 - 1) < 1e-4
    # return_types_test.cpp:14-28: j_l(k_{l,nmax-1} rmax) == 0 within 1e-15
    k = O.kln(10, 20, 5.4)
    for l in range(10):
        assert abs(O.lib().orc_sbessel_jl(l, k[l, 19] * 5.4)) < 1e-15


def test_kln_fixture(golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.k_ln"))
    k = O.kln(10, 10, 2500.0)
    np.testing.assert_allclose(k, ref, rtol=5e-7)  # fixture zeros are themselves off by 1.6e-7


def test_count_nl_config1():
    nl = O.count_nl(16, 40, 500.0, 0.1)
    assert nl[0] == 15 and nl[15] == 9 and int(((2 * np.arange(16) + 1) * nl).sum()) == 2744


# ---------------------------------------------------------------- direct path
def test_direct_fixture(gaussian_points, golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.c_ln"))
    _, c, _ = O.sfb_raw(gaussian_points, 10, 10, 2500.0)
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=2e-6)
    _, cc, _ = O.sfb_raw(gaussian_points, 10, 10, 2500.0, zeros_mode=O.ZEROS_CONVERGED)
    np.testing.assert_allclose(cc * CF ** 2, ref, rtol=1e-6)


def test_direct_fixture_float32_catalogue(golden_dir):
    # the C++ test reads the catalogue as float32 records (sfb_test.cpp:34)
    xyz = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.txt")).astype(np.float32).astype(np.float64)
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.c_ln"))
    _, c, _ = O.sfb_raw(workloads.to_spherical(xyz), 10, 10, 2500.0, zeros_mode=O.ZEROS_CONVERGED)
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=1e-6)


def test_direct_brute_force(golden_dir):
    g = np.load(os.path.join(golden_dir, "brute_force.npz"))
    lmax, nmax, rmax = int(g["lmax"]), int(g["nmax"]), float(g["rmax"])
    k, c, f = O.sfb_raw(g["points"], lmax, nmax, rmax, zeros_mode=O.ZEROS_CONVERGED, store_flmn=True)
    np.testing.assert_allclose(k, g["k_ln"], rtol=1e-14)
    np.testing.assert_allclose(c, g["c_ln"], rtol=1e-11)
    flat = np.concatenate([x.ravel() for x in f])
    assert np.max(np.abs(flat - g["f_lmn"])) < 1e-11 * np.max(np.abs(g["f_lmn"]))


def test_direct_single_point_closed_form():
    # one point: c_ln = (2/pi) V^2 k^2 j_l(k r)^2 / (4 pi) for any direction
    rmax, r = 300.0, 123.4
    for th, ph in [(0.3, 1.1), (2.9, 5.0), (1.5707963, 0.0)]:
        k, c, _ = O.sfb_raw(np.array([[r, th, ph]]), 8, 5, rmax)
        V = 4.0 / 3.0 * np.pi * rmax ** 3
        for l in range(8):
            j = O.sbessel_jl(l, k[l] * r)
            np.testing.assert_allclose(c[l], 2 / np.pi * V ** 2 * k[l] ** 2 * j ** 2 / (4 * np.pi), rtol=1e-11)


def _rotate(pts, R):
    xyz = workloads.to_cartesian(pts) @ R.T
    return workloads.to_spherical(xyz)


def test_direct_rotation_invariance():
    pts = workloads.uniform_ball(300, 200.0, seed=9)
    q, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))
    _, c1, f1 = O.sfb_raw(pts, 9, 4, 200.0, store_flmn=True)
    _, c2, _ = O.sfb_raw(_rotate(pts, q), 9, 4, 200.0)
    np.testing.assert_allclose(c2, c1, rtol=1e-10)
    # azimuthal rotation by a: f_lm -> f_lm e^{-i m a}
    a = 0.7
    rot = pts.copy()
    rot[:, 2] = np.mod(rot[:, 2] + a, 2 * np.pi)
    _, _, f3 = O.sfb_raw(rot, 9, 4, 200.0, store_flmn=True)
    for l in range(9):
        ph = np.exp(-1j * np.arange(l + 1) * a)[:, None]
        np.testing.assert_allclose(f3[l], f1[l] * ph, rtol=1e-9, atol=1e-9 * np.abs(f1[l]).max())


def test_direct_interpolated_table():
    pts = workloads.uniform_ball(500, 400.0, seed=2)
    _, ce, _ = O.sfb_raw(pts, 8, 6, 400.0)
    _, ci, _ = O.sfb_raw(pts, 8, 6, 400.0, interpolated=True)
    rel = np.abs(ci / ce - 1)
    assert 1e-9 < rel.max() < 1e-3  # linear table with 1000 nodes per zero
    y, zmax, dxi = O.interp_table(3, 6)
    assert zmax == O.bessel_zeros(3, 7)[6]
    assert abs(dxi * zmax - (6000 - 1)) < 1e-8
    assert y[0] == 0.0 and abs(y[-1]) < 1e-11  # ends at (the reference's estimate of) a zero of j_3


def test_direct_interpolated_domain_error():
    pts = np.array([[500.0, 1.0, 1.0]])  # r > rmax: k_ln r beyond z_{l,nmax}
    with pytest.raises(ValueError):
        O.sfb_raw(pts, 3, 4, 100.0, interpolated=True)


def test_direct_empty_catalogue():
    with pytest.raises(ValueError):
        O.sfb_raw(np.zeros((0, 3)), 3, 4, 100.0)


# ---------------------------------------------------------------- HEALPix geometry
def test_healpix_roundtrip_and_rings():
    for ns in [1, 2, 4, 8, 16]:
        npix = 12 * ns * ns
        for p in range(npix):
            th, ph = O.pix2ang_ring(ns, p)
            assert O.lib().orc_ang2pix_ring(ns, th, ph) == p
        start = 0
        for ring in range(1, 4 * ns):
            sp, rp, sh, z, ph0 = O.ring_info(ns, ring)
            assert sp == start
            th, ph = O.pix2ang_ring(ns, sp)
            assert abs(np.cos(th) - z) < 1e-14 and abs(ph - ph0) < 1e-14
            start += rp
        assert start == npix


def test_healpix_reference_known_answers():
    # point_types_tests.cpp:58-83
    p = O.lib().orc_ang2pix_ring(256, 1.2, 0.4)
    th, ph = O.pix2ang_ring(256, p)
    assert abs(th - 1.2) < 1e-2 * 1.2 and abs(ph - 0.4) < 1e-2 * 0.4
    with pytest.raises(ValueError):
        O.pixelize(np.zeros((1, 3)), 255)
    # test/python/point_test.py:44-50: two points on a pixel, one at the south pole
    off, rad = O.pixelize(np.array([[1, 0, 0], [1, 0, 0], [1, np.pi, 0]], dtype=float), 32)
    counts = np.diff(off)
    assert counts.sum() == 3 and np.count_nonzero(counts) == 2


def test_healpix_equal_area():
    pts = workloads.uniform_ball(120000, 1.0, seed=1)
    off, _ = O.pixelize(pts, 4)
    counts = np.diff(off)
    expect = 120000 / 192
    assert np.all(np.abs(counts - expect) < 6 * np.sqrt(expect))


# ---------------------------------------------------------------- HEALPix SFB
def test_healpix_fixture(gaussian_points, golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.pythonrev.c_ln"))
    off, rad = O.pixelize(gaussian_points, 64)
    _, c, f = O.sfb_reverse_fft(off, rad, 64, 10, 10, 2500.0, store_flmn=True)
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=2e-6)
    # no occupied ring is short enough to fold: equals the exact pixelised transform
    _, c2, f2 = O.sfb_reverse(off, rad, 64, 10, 10, 2500.0, store_flmn=True)
    np.testing.assert_allclose(c2, c, rtol=1e-11)


def test_healpix_alias_quirk_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "alias_quirk.npz"))
    ns, lmax, nmax, rmax = int(g["nside"]), int(g["lmax"]), int(g["nmax"]), float(g["rmax"])
    off, rad = O.pixelize(g["points"], ns)
    kw = dict(zeros_mode=O.ZEROS_CONVERGED)
    _, ca, _ = O.sfb_reverse_fft(off, rad, ns, lmax, nmax, rmax, alias_compat=True, **kw)
    _, cc, _ = O.sfb_reverse_fft(off, rad, ns, lmax, nmax, rmax, alias_compat=False, **kw)
    _, cr, _ = O.sfb_reverse(off, rad, ns, lmax, nmax, rmax, **kw)
    np.testing.assert_allclose(ca, g["c_as_written"], rtol=1e-10)
    np.testing.assert_allclose(cc, g["c_conjugated"], rtol=1e-10)
    np.testing.assert_allclose(cr, g["c_conjugated"], rtol=1e-10)
    # the quirk is real: it changes l >= 3 at nside 4 (polar rings of 4 and 8 pixels)
    assert np.all(np.abs(ca[:3] / cc[:3] - 1) < 1e-12)
    assert np.max(np.abs(ca[3:] / cc[3:] - 1)) > 1e-3


# ---------------------------------------------------------------- output format
def test_savetxt_format(tmp_path):
    a = np.array([[1.115581300132204346e+12, -0.0], [5.0e-300, 3.0]])
    p = str(tmp_path / "x.c_ln")
    O.savetxt(p, a)
    lines = open(p).read().splitlines()
    assert lines[0] == "1.115581300132204346e+12 -0.000000000000000000e+00"
    # std::ostream with precision(18) + scientific prints what printf("%.18e") prints
    assert lines[1] == "%.18e %.18e" % (5.0e-300, 3.0)
    np.testing.assert_array_equal(np.loadtxt(p), a)


# ---------------------------------------------------------------- double spherical-Bessel integral
def _lommel(l, R, a, b):
    """int_0^R r^2 j_l(a r) j_l(b r) dr, a != b (Lommel's integral)."""
    from scipy.special import spherical_jn as J
    jm = (lambda x: np.cos(x) / x) if l == 0 else (lambda x: J(l - 1, x))
    return R ** 2 * (b * J(l, a * R) * jm(b * R) - a * jm(a * R) * J(l, b * R)) / (a * a - b * b)


def _diag(l, R, k):
    """int_0^R r^2 j_l(k r)^2 dr = R^3/2 [j_l(kR)^2 - j_{l-1}(kR) j_{l+1}(kR)]."""
    from scipy.special import spherical_jn as J
    x = k * R
    jm = np.cos(x) / x if l == 0 else J(l - 1, x)
    return R ** 3 / 2 * (J(l, x) ** 2 - jm * J(l + 1, x))


def test_sbessel_integral_reference_known_values():
    # Mathematica values of test/besseltools_tests/sbessel_integrator.cpp:15-19
    # (test/python/besseltools_test.py:27-31), w = 1
    assert abs(O.sbessel_integral(0, 100, 1, 1) - 1.56582) < 1e-5
    assert abs(O.sbessel_integral(0, 100, 1, 2) - 0.785393) < 1e-5
    assert abs(O.sbessel_integral(4, 100, 1, 4.2) - 0.000133933) < 1e-9


def test_sbessel_integral_symmetric_in_k():
    # sbessel_integrator.cpp (test) :8-13, w = exp(-(r/40)^2) r^2
    a = O.sbessel_integral(6, 100, 2.3, 4.5, power=2, gauss=40)
    b = O.sbessel_integral(6, 100, 4.5, 2.3, power=2, gauss=40)
    assert abs(a - b) <= 1e-12 * abs(a)


@pytest.mark.parametrize("l,k", [(0, 0.3), (1, 0.05), (5, 0.2), (40, 0.11), (255, 0.3), (255, 0.5)])
def test_sbessel_integral_diagonal_closed_form(l, k):
    R = 1000.0
    v = O.sbessel_integral(l, R, k, k, power=2)
    assert abs(v - _diag(l, R, k)) <= 1e-10 * R / k ** 2


@pytest.mark.parametrize("l,a,b", [(0, 0.013, 0.037), (5, 0.2, 0.21), (17, 0.1, 0.35), (255, 0.27, 0.5)])
def test_sbessel_integral_lommel(l, a, b):
    R = 1000.0
    v = O.sbessel_integral(l, R, a, b, power=2)
    assert abs(v - _lommel(l, R, a, b)) <= 1e-10 * R / (a * b)


def test_sbessel_integral_linear_table_is_exact():
    # a linear w is reproduced exactly by the interpolating table
    R, l, k1, k2 = 300.0, 3, 0.07, 0.11
    x = np.linspace(0.0, 400.0, 801)
    table = (2.0 + 0.5 * x, 0.0, 400.0)
    v = O.sbessel_integral(l, R, k1, k2, table=table)
    ref = 2.0 * O.sbessel_integral(l, R, k1, k2) + 0.5 * O.sbessel_integral(l, R, k1, k2, power=1)
    assert abs(v - ref) <= 1e-11 * R / (k1 * k2)
    with pytest.raises(ValueError):  # the table must cover [0, R)
        O.sbessel_integral(l, R, k1, k2, table=(np.ones(11), 0.0, 200.0))


def test_sbessel_integral_stopping_rule_keeps_the_value():
    # a narrow Gaussian window: outer ranges past r ~ 300 change the sum by less
    # than 1e-10 and the loop stops there; scipy's quad on [0, 400] agrees
    from scipy.integrate import quad
    from scipy.special import spherical_jn as J
    l, k, g, R = 3, 0.1, 50.0, 1000.0
    v = O.sbessel_integral(l, R, k, k, power=2, gauss=g)
    ref = quad(lambda r: r * r * np.exp(-(r / g) ** 2) * J(l, k * r) ** 2, 0, 400, limit=400, epsabs=0,
               epsrel=1e-13)[0]
    assert abs(v - ref) <= 1e-9 * ref


def test_interpolated_lookup_vs_numpy_interp():
    # one point: c_ln = (2/pi) V^2 k^2 j(k r)^2 / (4 pi) whatever the angles, so the oracle's
    # linear-table value j(k r) is read back and compared with numpy.interp over nodes
    # x_i = i z_max / (P - 1), P = 1000 nmax, z_max = z_{l,nmax}, valued by scipy's j_l
    # (SBesselInterpolator.cpp:10-25, FunctionInterpolator.cpp:12-27)
    from scipy.special import spherical_jn
    lmax, nmax, rmax = 6, 5, 300.0
    V = 4.0 / 3.0 * np.pi * rmax ** 3
    for r in (0.37, 12.5, 151.3, 299.9):
        pts = np.array([[r, 1.1, 0.4]])
        k, c, _ = O.sfb_raw(pts, lmax, nmax, rmax, interpolated=True)
        jtab = np.sqrt(c * 4 * np.pi / (2 / np.pi * V ** 2 * k ** 2))
        for l in range(lmax):
            zmax = O.bessel_zeros(l, nmax + 1)[nmax]
            x = np.linspace(0.0, zmax, 1000 * nmax)
            x = np.arange(1000 * nmax) * (zmax / (1000 * nmax - 1))
            y = spherical_jn(l, x)
            expect = np.abs(np.interp(k[l] * r, x, y))
            np.testing.assert_allclose(jtab[l], expect, rtol=1e-10, atol=1e-13)
            # and the table is not the exact function: the interpolation error is visible
            exact = np.abs(spherical_jn(l, k[l] * r))
            assert np.abs(expect - exact).max() > 1e-12 or r < 1.0
