"""The Python drop-in API (catana's names) on the GPU: the cases of the
reference's test/python/decomposition_test.py and point_test.py, plus parity
of every entry point with the oracle, the output formats and error types."""
import os

import numpy as np
import pytest

import catana_amd as catana
from catana_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CF = 5535184878.03 / (2500 ** 3 * 4 / 3 * np.pi)  # decomposition_test.py:15


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def gaussian(golden_dir):
    xyz = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.txt"))
    return {
        "pc": catana.PointContainer(xyz, "cartesian"),
        "c_ln": np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.c_ln")),
        "c_ln_rev": np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.pythonrev.c_ln")),
        "k_ln": np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.k_ln")),
    }


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


def test_raw(gaussian):
    print("case: sfb_decomposition(PointContainer) gaussian fixture", flush=True)
    pc = gaussian["pc"]
    assert np.array(pc).shape == (445, 3)
    kclkk = catana.sfb_decomposition(pc, 10, 10, 2500, False, False)
    assert np.allclose(kclkk.k_ln, gaussian["k_ln"])
    assert np.allclose(kclkk.c_ln * CF ** 2, gaussian["c_ln"], rtol=1e-2)
    np.testing.assert_allclose(kclkk.c_ln * CF ** 2, gaussian["c_ln"], rtol=2e-6)
    assert kclkk.f_lmn == []


def test_rev(gaussian):
    print("case: sfb_decomposition(PixelizedPointContainer) gaussian fixture nside 64", flush=True)
    ppc = catana.PixelizedPointContainer(64, gaussian["pc"])
    assert ppc.get_npoints() == 445
    kclkk = catana.sfb_decomposition(ppc, 10, 10, 2500, False, False)
    assert np.allclose(kclkk.k_ln, gaussian["k_ln"])
    assert np.allclose(kclkk.c_ln * CF ** 2, gaussian["c_ln_rev"], rtol=1e-2)
    np.testing.assert_allclose(kclkk.c_ln * CF ** 2, gaussian["c_ln_rev"], rtol=2e-6)


def test_pc_constructor_countmap():
    # point_test.py: TestPixelizedPointContainer.test_pc_constructor
    pc = catana.PointContainer(np.array([[1, 0, 0], [1, 0, 0], [1, np.pi, 0]]))
    cm = catana.PixelizedPointContainer(32, pc).get_countmap()
    assert cm.shape == (32 ** 2 * 12,)
    assert np.sum(cm) == 3
    assert np.count_nonzero(cm) == 2
    # point_types_tests.cpp:87-109
    pc = catana.PointContainer()
    for xyz in [(-1, 2, 3), (-4, 1, 0), (0, 1, 2), (1, 4, 0), (1, 2, 3)]:
        pc.add_point(catana.Point(*xyz))
    ppc = catana.PixelizedPointContainer(16, pc)
    assert ppc.get_npoints() == 5 and ppc.get_countmap().shape == (16 * 16 * 12,) and ppc.get_countmap().sum() == 5


def test_countmap_matches_oracle():
    pts = workloads.uniform_ball(30000, 50.0, seed=5)
    ppc = catana.PixelizedPointContainer(8, catana.PointContainer(pts))
    off, _ = O.pixelize(pts, 8)
    np.testing.assert_array_equal(ppc.get_countmap(), np.diff(off))
    # adding points one by one on the host bins them the same way
    ppc2 = catana.PixelizedPointContainer(8)
    pc = catana.PointContainer(pts[:2000])
    for i in range(2000):
        ppc2.add_point(pc[i])
    np.testing.assert_array_equal(ppc2.get_countmap(), np.diff(O.pixelize(pts[:2000], 8)[0]))


def test_kclkk_and_savetxt(tmp_path):
    k = catana.KClkk(12, 7, 300.0)
    np.testing.assert_allclose(k.k_ln, O.kln(12, 7, 300.0), rtol=1e-15, atol=0)
    assert np.all(k.c_ln == 0) and k.c_ln.shape == (12, 7)
    np.testing.assert_array_equal(catana.compute_kln(12, 7, 300.0), k.k_ln)
    # byte-exact text format (return_types.cpp:39-64) against the oracle's writer
    pts = workloads.uniform_ball(500, 300.0, seed=2)
    r = catana.sfb_decomposition(catana.PointContainer(pts), 12, 7, 300.0, False, False)
    r.savetxt(str(tmp_path / "a"))
    O.savetxt(str(tmp_path / "b.k_ln"), r.k_ln)
    O.savetxt(str(tmp_path / "b.c_ln"), r.c_ln)
    for ext in ("k_ln", "c_ln"):
        assert (tmp_path / f"a.{ext}").read_bytes() == (tmp_path / f"b.{ext}").read_bytes()
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "a.c_ln"), r.c_ln)


def test_store_flmn_layout_and_values():
    pts = workloads.uniform_ball(3000, 400.0, seed=4)
    pc = catana.PointContainer(pts)
    r = catana._sfb_raw(pc, 9, 6, 400.0, True, False, True, False)
    assert len(r.f_lmn) == 9
    for l, f in enumerate(r.f_lmn):
        assert f.shape == (l + 1, 6) and f.dtype == np.complex128
    ko, co, fo = O.sfb_raw(pts, 9, 6, 400.0, store_flmn=True)
    np.testing.assert_allclose(r.c_ln, co, rtol=1e-10)
    assert _flmn_err(r.f_lmn, fo) < 1e-10
    # public wrapper: N > 1000 selects the reference's interpolation table
    r2 = catana.sfb_decomposition(pc, 9, 6, 400.0, True, False)
    _, co2, fo2 = O.sfb_raw(pts, 9, 6, 400.0, interpolated=True, store_flmn=True)
    np.testing.assert_allclose(r2.c_ln, co2, rtol=1e-10)
    assert _flmn_err(r2.f_lmn, fo2) < 1e-10


def test_reverse_paths_vs_oracle():
    pts = workloads.uniform_ball(20000, 200.0, seed=8)
    ppc = catana.PixelizedPointContainer(4, catana.PointContainer(pts))
    off, rad = O.pixelize(pts, 4)
    for fn, ofn in ((catana._sfb_reverse_fft, O.sfb_reverse_fft), (catana._sfb_reverse, O.sfb_reverse)):
        r = fn(ppc, 12, 3, 200.0, True, False, True, False)
        _, co, fo = ofn(off, rad, 4, 12, 3, 200.0, store_flmn=True)
        np.testing.assert_allclose(r.c_ln, co, rtol=1e-10)
        assert _flmn_err(r.f_lmn, fo) < 1e-10


def test_k_max_entry_points():
    pts = workloads.uniform_ball(20000, 500.0, seed=9)
    pc = catana.PointContainer(pts)
    lmax, k_max, rmax, nmax, vol = 16, 0.1, 500.0, 16, 3.1e8
    nl = O.count_nl(lmax, nmax, rmax, k_max)
    r64 = catana.sfb_decomposition_reverse(pc, lmax, k_max, rmax, nmax, vol, True, False, "fp64")
    _, co, fo = O.sfb_raw(pts, lmax, nmax, rmax, nl=nl, volume=vol, store_flmn=True)
    np.testing.assert_allclose(r64.c_ln, co, rtol=1e-10)
    assert _flmn_err(r64.f_lmn, fo) < 1e-10
    for l in range(lmax):
        assert np.all(r64.c_ln[l, nl[l]:] == 0)
    rmx = catana.sfb_decomposition_reverse(pc, lmax, k_max, rmax, nmax, vol, True, False)  # mixed default
    nz = r64.c_ln > 0
    assert np.max(np.abs(rmx.c_ln[nz] / r64.c_ln[nz] - 1)) < 1e-6
    assert np.all(rmx.c_ln[~nz] == 0)
    assert _flmn_err(rmx.f_lmn, r64.f_lmn) < 1e-6
    with pytest.raises(ValueError):
        catana.sfb_decomposition_reverse(pc, lmax, k_max, rmax, nmax, vol, False, False, "half")
    # HEALPix variant from points and from a pixelised container
    h1 = catana.sfb_decomposition_healpix(pc, 8, lmax, k_max, rmax, nmax, vol, True, False)
    h2 = catana.sfb_decomposition_healpix(catana.PixelizedPointContainer(8, pc), lmax, k_max, rmax, nmax, vol,
                                          True, False)
    off, rad = O.pixelize(pts, 8)
    _, co, fo = O.sfb_reverse_fft(off, rad, 8, lmax, nmax, rmax, nl=nl, volume=vol, store_flmn=True)
    np.testing.assert_array_equal(h1.c_ln, h2.c_ln)
    np.testing.assert_allclose(h1.c_ln, co, rtol=1e-10)
    assert _flmn_err(h1.f_lmn, fo) < 1e-10
    # the ring path through the same entry point (precision 'mixed'); l < 2 run in FP64,
    # the stated tolerances are 1e-6 per-l f_lmn and 5e-6 C_l below l = 8
    hm = catana.sfb_decomposition_healpix(pc, 8, lmax, k_max, rmax, nmax, vol, True, False, "mixed")
    nz = h1.c_ln > 0
    assert np.all(hm.c_ln[~nz] == 0)
    assert np.max(np.abs(hm.c_ln[nz] / h1.c_ln[nz] - 1)) < 5e-6
    assert _flmn_err(hm.f_lmn, h1.f_lmn) < 1e-6


def test_error_types():
    pts = workloads.uniform_ball(2000, 100.0, seed=1)
    pts[7, 0] = 250.0  # beyond the interpolation range of the reference's table
    pc = catana.PointContainer(pts)
    with pytest.raises(ValueError, match="range"):
        catana.sfb_decomposition(pc, 6, 4, 100.0, False, False)
    # the exact-Bessel entry accepts it
    r = catana._sfb_raw(pc, 6, 4, 100.0, False, False, True, False)
    assert np.all(np.isfinite(r.c_ln))
    with pytest.raises(ValueError, match="power of 2"):
        catana.PixelizedPointContainer(12, pc)


def test_batched_api_matches_single_calls():
    specs = [(workloads.uniform_ball(5000, 400.0, seed=61), 20, 0.1, 400.0, 16, 0.0),
             (workloads.clustered_ball(300, 900.0, mean_children=30, sigma=5.0, seed=62), 36, 0.15, 900.0, 50, 4e9),
             (workloads.uniform_ball(700, 100.0, seed=63), 8, 0.0, 100.0, 5, 0.0)]
    cats = [(catana.PointContainer(p), L, k, r, n, v) for p, L, k, r, n, v in specs]
    print("case: batched Python API, 3 catalogues", flush=True)
    for prec, tol in (("fp64", 1e-10), ("mixed", 1e-6)):
        res = catana.sfb_decomposition_reverse_batched(cats, store_flmn=True, verbose=False, precision=prec)
        assert len(res) == 3
        for (pc, L, k, r, n, v), b in zip(cats, res):
            s = catana.sfb_decomposition_reverse(pc, L, k, r, n, v, True, False, prec)
            np.testing.assert_array_equal(b.k_ln, s.k_ln)
            nz = s.c_ln != 0
            assert np.all((b.c_ln != 0) == nz)
            assert np.max(np.abs(b.c_ln[nz] / s.c_ln[nz] - 1)) < tol
            assert len(b.f_lmn) == L and b.f_lmn[-1].shape == (L, n)
            assert _flmn_err(b.f_lmn, s.f_lmn) < tol
    with pytest.raises(ValueError):
        catana.sfb_decomposition_reverse_batched([(cats[0][0], 4)], verbose=False)
