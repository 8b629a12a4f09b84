"""Ingestion before the path, host side (no device work): the oracle's
restatement pinned by the reference's own fixtures and known values, and the
C++ sources and sinks against it (test/io_tests/*)."""
import os

import numpy as np
import pytest

from oracle import ingest as I

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# test/io_tests/read_gadget_tests.cpp:13-17 (radii of the mock box times 0.7)
POS = [0, 10, 100, 511]
RAD = np.array([11.43271463, 7.47124219, 5.63280499, 8.62107958]) * 0.7
THETA = [0.90536372, 1.8781445, 1.42947137, 0.87425136]
PHI = [4.04660937, 2.79354124, 2.96859039, 5.80961737]


def test_splitmix64_known_value():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 0: first output
    assert int(I.splitmix64(np.uint64(0x9E3779B97F4A7C15))) == 0xE220A8397B1DCDAF


def test_filter_uniform_is_uniform():
    u = I.filter_uniform(1234, 0, np.arange(200000))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    h, _ = np.histogram(u, bins=20, range=(0, 1))
    assert np.all(np.abs(h / 10000.0 - 1) < 0.05)
    # other slots and seeds draw other numbers
    assert not np.array_equal(u[:100], I.filter_uniform(1234, 1, np.arange(100)))
    assert not np.array_equal(u[:100], I.filter_uniform(99, 0, np.arange(100)))


@pytest.mark.parametrize("fname", ["mock_data_z0p000.0", "mock_data_z0p000.1"])
def test_oracle_gadget_known_values(fname):
    a = I.read_gadget(os.path.join(GOLD, fname))
    assert a.shape == (512, 3)
    np.testing.assert_allclose(a[POS, 0], RAD, rtol=1e-5)
    np.testing.assert_allclose(a[POS, 1], THETA, rtol=1e-5)
    np.testing.assert_allclose(a[POS, 2], PHI, rtol=1e-5)


def test_gadget_source_matches_oracle():
    import catana_amd as catana
    for fname in ("mock_data_z0p000.0", "mock_data_z0p000.1"):
        src = catana.io.GadgetSource(os.path.join(GOLD, fname), False)
        a = np.array(src.get_point_container())
        np.testing.assert_allclose(a, I.read_gadget(os.path.join(GOLD, fname)), rtol=1e-15, atol=1e-15)
        src.reset()
        assert np.array_equal(np.array(src.get_point_container()), a)


def test_text_sources_and_sinks(tmp_path):
    import catana_amd as catana
    src = catana.io.CartesianTextSource(os.path.join(GOLD, "mock_data.txt"))
    a = np.array(src.get_point_container())
    assert a.shape == (20, 3)
    assert a[0, 0] == pytest.approx(8.002900041)  # readwrite_text_tests.cpp:26
    np.testing.assert_allclose(a, I.cartesian_to_points(np.loadtxt(os.path.join(GOLD, "mock_data.txt"))),
                               rtol=1e-15, atol=1e-15)
    # shift moves the origin on every axis
    b = np.array(catana.io.CartesianTextSource(os.path.join(GOLD, "mock_data.txt"), 1.5).get_point_container())
    np.testing.assert_allclose(b, I.cartesian_to_points(np.loadtxt(os.path.join(GOLD, "mock_data.txt")), 1.5),
                               rtol=1e-15, atol=1e-15)
    pc = catana.PointContainer(a, "spherical")
    for sink_cls, src_cls, tol in [(catana.io.CartesianTextSink, catana.io.CartesianTextSource, 1e-14),
                                   (catana.io.SphericalTextSink, catana.io.SphericalTextSource, 0.0),
                                   (catana.io.SphericalTextSink_3dex, catana.io.SphericalTextSource_3dex, 1e-11)]:
        fn = str(tmp_path / (sink_cls.__name__ + ".txt"))
        sink = sink_cls(fn, False)
        assert sink.write(pc) == 20
        sink.close()
        back = np.array(src_cls(fn).get_point_container())
        np.testing.assert_allclose(back, a, rtol=tol, atol=tol)
    # the 3dex format of the reference: " %18.14F %18.14F %18.12F" (phi, theta, r)
    line = open(str(tmp_path / "SphericalTextSink_3dex.txt")).readline()
    assert line == " %18.14F %18.14F %18.12F\n" % (a[0, 2], a[0, 1], a[0, 0])


def test_raw_binary_round_trip_and_append(tmp_path):
    # readwrite_rawdata_tests.cpp:9-57 (spherical float records)
    import catana_amd as catana
    a = np.array([[1.0, 1.0, 3.0], [2.0, 0.1, 0.0], [3.0, 0.2, 1.0], [100.0, 0.3, 4.0]])
    pc = catana.PointContainer(a, "spherical")
    fn = str(tmp_path / "raw.bin")
    s = catana.io.RawBinarySink(fn, False, False)
    assert s.write(pc) == 4
    s.close()
    assert s.write(pc) == -1
    s = catana.io.RawBinarySink(fn, False, True)
    assert s.write(pc) == 4
    s.close()
    back = np.array(catana.io.RawBinarySource(fn, False).get_point_container())
    assert back.shape == (8, 3)
    np.testing.assert_allclose(back, np.concatenate([a, a]).astype(np.float32), rtol=0, atol=0)


def test_angular_mask_construction_errors():
    import catana_amd as catana
    with pytest.raises(RuntimeError):  # filter_tests.cpp:118-121
        catana.io.AngularMaskFilter(np.zeros(200, dtype=np.float32))
    with pytest.raises(RuntimeError, match="FITS"):
        catana.io.AngularMaskFilter(os.path.join(GOLD, "mock_data.txt"))
