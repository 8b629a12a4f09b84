"""Host side of the Python API (catana's basictypes), no device work:
the cases of the reference's test/python/point_test.py and
test/point_types_tests/point_types_tests.cpp:11-27, 65-90."""
import numpy as np
import pytest

import catana_amd as catana
from oracle import oracle as O


class TestPoint:
    def test_default_constructor(self):
        p = catana.Point()
        assert p.cartesian() == (0.0, 0.0, 0.0)
        assert p.spherical() == (0.0, 0.0, 0.0)
        assert np.allclose(p, np.zeros(3))

    def test_value_constructor(self):
        p = catana.Point(0, 0, 1)
        assert p.spherical() == (1, 0, 0)
        assert p.cartesian() == (0, 0, 1)

    def test_cartesian_values(self):
        # point_types_tests.cpp:11-27
        for xyz, sph in [((0, 0, 1), (1, 0, 0)),
                         ((1, 1, 1), (1.73205081, 9.55316618e-01, 7.85398163e-01)),
                         ((-380, -800, 200), (9.07964757e+02, 1.34870213e+00, 4.26894064e+00))]:
            np.testing.assert_allclose(catana.Point(*xyz).spherical(), sph, rtol=1e-8, atol=1e-12)

    def test_phi_range(self):
        rng = np.random.default_rng(3)
        for x, y, z in rng.normal(size=(200, 3)):
            r, t, p = catana.Point(x, y, z).spherical()
            assert 0 <= p < 2 * np.pi and 0 <= t <= np.pi
            np.testing.assert_allclose(catana.Point(x, y, z).cartesian(), (x, y, z), rtol=1e-12, atol=1e-14)


class TestPointContainer:
    def test_default_constructor(self):
        assert np.array(catana.PointContainer()).shape == (0, 3)

    def test_add_point(self):
        pc = catana.PointContainer()
        pc.add_point(catana.Point())
        assert np.array(pc).shape == (1, 3)
        assert np.allclose(pc, np.zeros((1, 3)))

    def test_numpy_constructor(self):
        data = np.arange(9).reshape(3, 3)
        pc = catana.PointContainer(data)
        assert np.allclose(pc, data)

    def test_cartesian_constructor(self):
        xyz = np.random.default_rng(1).normal(size=(50, 3))
        pc = np.array(catana.PointContainer(xyz, "cartesian"))
        np.testing.assert_allclose(pc[:, 0], np.linalg.norm(xyz, axis=1), rtol=1e-14)
        np.testing.assert_allclose(np.cos(pc[:, 1]) * pc[:, 0], xyz[:, 2], rtol=1e-12, atol=1e-14)

    def test_getitem(self):
        pc = catana.PointContainer(np.ones((2, 3)))
        assert np.allclose(pc[0], pc[1])
        assert np.allclose(pc[0], [1, 1, 1])
        with pytest.raises(IndexError):
            pc[2]

    def test_bad_arrays(self):
        with pytest.raises(RuntimeError, match="dimensions"):
            catana.PointContainer(np.ones(3))
        with pytest.raises(RuntimeError, match="columns"):
            catana.PointContainer(np.ones((3, 4)))
        with pytest.raises(RuntimeError, match="RowMajor"):
            catana.PointContainer(np.asfortranarray(np.ones((3, 3))))
        with pytest.raises(RuntimeError, match="coord_type"):
            catana.PointContainer(np.ones((3, 3)), "polar")


class TestPixelizedPointContainer:
    def test_default_constructor(self):
        with pytest.raises(TypeError):
            catana.PixelizedPointContainer()

    def test_empty_constructor(self):
        ppc = catana.PixelizedPointContainer(4)
        assert np.all(ppc.get_countmap() == np.zeros(4 ** 2 * 12))
        assert len(catana.PixelizedPointContainer(256)) == 12 * 256 * 256

    def test_non_power_of_two(self):
        with pytest.raises(ValueError):
            catana.PixelizedPointContainer(255)

    def test_add_point(self):
        # point_types_tests.cpp:65-84
        ppc = catana.PixelizedPointContainer(256)
        ppc.add_point(catana.PointContainer(np.array([[10.0, 1.2, 0.4]]))[0])
        cm = ppc.get_countmap()
        assert ppc.get_npoints() == 1 and cm.sum() == 1
        q = int(np.flatnonzero(cm)[0])
        assert q == int(O.ang2pix_ring(256, np.array([1.2]), np.array([0.4]))[0])
        t, p = O.pix2ang_ring(256, q)
        assert abs(t - 1.2) < 1e-2 * 1.2 and abs(p - 0.4) < 1e-2 * 0.4


# ---- besseltools and FunctionInterpolator (host behaviour only) ----

def test_spherical_bessel_zeros_api():
    # test/python/besseltools_test.py:6-23
    import catana_amd as catana
    with pytest.raises(TypeError):
        catana.besseltools.SphericalBesselZeros()
    z = catana.besseltools.SphericalBesselZeros(0)
    assert z[0] == pytest.approx(np.pi)
    assert z[1] == pytest.approx(6.283185307179586)
    assert z[2] == pytest.approx(9.42477796076938)
    assert catana.besseltools.SphericalBesselZeros(5)[9] == pytest.approx(38.8836)
    z2 = catana.besseltools.SphericalBesselZeros(2)
    z2.compute_up_to(100)
    assert z2[30] > 100 > z2[29] or z2[29] > 100  # zeros extend past 100


def test_function_interpolator_api():
    # test/python/besseltools_test.py:39-52
    import catana_amd as catana
    w = lambda r: np.exp(-(r / 40) ** 2) * r ** 2  # noqa: E731
    fi = catana.FunctionInterpolator(w, 100000, 0, 100)
    assert fi(0) == w(0)
    assert fi(99.9) == pytest.approx(w(99.9))
    assert fi(20) == pytest.approx(fi(20))
    for bad in (100, -1, 120):
        with pytest.raises(ValueError):
            fi(bad)
    x = np.array([1.0, 2.5, 50.0])
    np.testing.assert_allclose(fi(x), w(x), rtol=1e-8)
    # linear interpolation between samples: exact for a linear function
    fl = catana.FunctionInterpolator(lambda r: 3.0 - 2.0 * r, 11, 0.0, 10.0)
    assert fl(3.7) == pytest.approx(3.0 - 7.4, rel=1e-15)
