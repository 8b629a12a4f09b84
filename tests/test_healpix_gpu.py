"""Parity of the HEALPix path (binning + pixelised SFB, FP64) with the oracle,
the reference's fixture and the numpy.fft golden of the alias quirk."""
import os

import numpy as np
import pytest
import torch  # noqa: F401

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

CF = 5535184878.03 / (2500 ** 3 * 4 / 3 * np.pi)


def _pix(handle, pts, nside):
    d = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    off, rad = capi.pixelize_torch(handle, d, nside)
    torch.cuda.synchronize()
    return off, rad


def _hp(handle, off, rad, nside, lmax, nmax, rmax, **kw):
    k, c, f = capi.sfb_healpix_torch(handle, off, rad, nside, lmax, nmax, rmax, store_flmn=True, **kw)
    torch.cuda.synchronize()
    return k.cpu().numpy(), c.cpu().numpy(), capi.split_flmn(f.cpu().numpy(), lmax, nmax)


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


def test_pixelize_matches_oracle(handle):
    pts = workloads.uniform_ball(50000, 100.0, seed=31)
    print("case: pixelize N=50000 nside=16", flush=True)
    off, rad = _pix(handle, pts, 16)
    oo, ro = O.pixelize(pts, 16)
    np.testing.assert_array_equal(off.cpu().numpy(), oo)
    np.testing.assert_array_equal(rad.cpu().numpy(), ro)  # stable: input order kept within a pixel


def test_pixelize_rejects_non_power_of_two(handle):
    pts = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        capi.pixelize_torch(handle, pts, 255)


def test_healpix_fixture(handle, gaussian_points, golden_dir):
    print("case: gaussian fixture nside=64 lmax=nmax=10", flush=True)
    off, rad = _pix(handle, gaussian_points, 64)
    k, c, f = _hp(handle, off, rad, 64, 10, 10, 2500.0, bessel=capi.BESSEL_EXACT)
    oo, ro = O.pixelize(gaussian_points, 64)
    ko, co, fo = O.sfb_reverse_fft(oo, ro, 64, 10, 10, 2500.0, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.pythonrev.c_ln"))
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=2e-6)


@pytest.mark.parametrize("alias", [True, False])
def test_healpix_alias_quirk_golden(handle, golden_dir, alias):
    g = np.load(os.path.join(golden_dir, "alias_quirk.npz"))
    ns, lmax, nmax, rmax = int(g["nside"]), int(g["lmax"]), int(g["nmax"]), float(g["rmax"])
    print(f"case: alias quirk golden nside={ns} alias_compat={alias}", flush=True)
    off, rad = _pix(handle, g["points"], ns)
    _, c, _ = _hp(handle, off, rad, ns, lmax, nmax, rmax, bessel=capi.BESSEL_EXACT,
                  zeros=capi.ZEROS_CONVERGED, alias_compat=alias)
    np.testing.assert_allclose(c, g["c_as_written" if alias else "c_conjugated"], rtol=1e-10)


def test_healpix_uniform_vs_oracle_table_mode(handle):
    pts = workloads.uniform_ball(20000, 300.0, seed=32)
    print("case: uniform N=20000 nside=8 lmax=20 nmax=6 reference table, alias on", flush=True)
    off, rad = _pix(handle, pts, 8)
    k, c, f = _hp(handle, off, rad, 8, 20, 6, 300.0, bessel=capi.BESSEL_AUTO)
    oo, ro = O.pixelize(pts, 8)
    ko, co, fo = O.sfb_reverse_fft(oo, ro, 8, 20, 6, 300.0, interpolated=True, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10
    # without the quirk the path equals the exact pixelised transform
    _, c2, _ = _hp(handle, off, rad, 8, 20, 6, 300.0, bessel=capi.BESSEL_AUTO, alias_compat=False)
    _, cr, _ = O.sfb_reverse(oo, ro, 8, 20, 6, 300.0, interpolated=True)
    np.testing.assert_allclose(c2, cr, rtol=1e-10)
