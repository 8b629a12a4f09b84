"""The ring-structured HEALPix path (sfbhip_healpix with precision MIXED): FP32
j_l-weighted maps, ring Legendre table, split-FP16 ring contraction.

Stated tolerance, against the FP64 HEALPix path with exact j_l (the same
pixelised transform, sfb_reverse_fft.cpp:127-153 / :174-209), with and without
the reference's unconjugated fold (alias_compat):
  * per-l relative f_lmn norm error <= 1e-6;
  * relative C_l error <= 1e-6 for l >= 8 and <= 5e-6 for mixed_fp64_below <= l < 8
    (C_l of few m, closest to their noise: a pixel sum of coherent j_l values
    is rounded to 22 bits as one operand);
  * l < mixed_fp64_below (default 2) run in the FP64 pixel kernel (<= 1e-10);
    with mixed_fp64_below = 8 every C_l is within 1e-6.
The FP64 path's default for N > 1000 is the reference's linear j_l table, whose
own error (~1e-6 of the amplitude) then dominates the difference."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-6        # f_lmn per l; C_l for l >= 8
TOL_LOW_L = 5e-6  # C_l for mixed_fp64_below <= l < 8
CF = 5535184878.03 / (2500 ** 3 * 4 / 3 * np.pi)  # test/python/decomposition_test.py:15


def _pix(handle, pts, nside):
    d = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    off, rad = capi.pixelize_torch(handle, d, nside)
    torch.cuda.synchronize()
    return off, rad


def _hp(handle, off, rad, nside, lmax, nmax, rmax, **kw):
    k, c, f = capi.sfb_healpix_torch(handle, off, rad, nside, lmax, nmax, rmax, store_flmn=True, **kw)
    torch.cuda.synchronize()
    return k.cpu().numpy(), c.cpu().numpy(), capi.split_flmn(f.cpu().numpy(), lmax, nmax)


def _cmp(cr, c64, fr, f64, what, low_l_tol=TOL_LOW_L):
    assert np.all((cr == 0) == (c64 == 0)), what
    m = c64 != 0
    full = np.where(m, np.abs(cr / np.where(m, c64, 1) - 1), 0)
    worst = np.unravel_index(np.argmax(full), full.shape)
    ferr = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(fr, f64) if np.linalg.norm(b) > 0]
    lo, hi = full[:8].max(), full[8:].max() if full.shape[0] > 8 else 0.0
    print(f"{what}: max rel C_l {full.max():.3e} at (l, n) = {tuple(int(x) for x in worst)} (l < 8: {lo:.3e}, "
          f"l >= 8: {hi:.3e}) median {np.median(full[m]):.3e}, max per-l f_lmn {max(ferr):.3e}", flush=True)
    assert hi <= TOL
    assert lo <= low_l_tol
    assert max(ferr) <= TOL
    return full.max()


# nside 128 and 256 run the equatorial tiles through healpix_ring_eq_kernel
@pytest.mark.parametrize("nside,lmax,alias", [(4, 20, True), (8, 30, True), (8, 30, False), (32, 64, True),
                                              (64, 40, False), (128, 90, True), (256, 48, False)])
def test_ring_vs_fp64_healpix(handle, nside, lmax, alias):
    pts = workloads.radial_selection(60000, 800.0, r0=300.0, seed=50 + nside)
    print(f"case: ring path nside={nside} lmax={lmax} alias={alias}", flush=True)
    off, rad = _pix(handle, pts, nside)
    kw = dict(k_max=0.1, bessel=capi.BESSEL_EXACT, alias_compat=alias)
    _, c64, f64 = _hp(handle, off, rad, nside, lmax, 40, 800.0, **kw)
    _, cr, fr = _hp(handle, off, rad, nside, lmax, 40, 800.0, precision=capi.PREC_MIXED, **kw)
    _cmp(cr, c64, fr, f64, f"nside {nside} lmax {lmax} alias {alias}")


def test_ring_vs_oracle_and_fixture(handle, gaussian_points, golden_dir):
    # the reference's fixture (nside 64, lmax = nmax = 10, rmax 2500) and the oracle's
    # restatement of _sfb_reverse_fft with exact j_l
    print("case: ring path, gaussian fixture nside 64", flush=True)
    off, rad = _pix(handle, gaussian_points, 64)
    k, c, f = _hp(handle, off, rad, 64, 10, 10, 2500.0, precision=capi.PREC_MIXED, bessel=capi.BESSEL_EXACT)
    oo, ro = O.pixelize(gaussian_points, 64)
    _, co, fo = O.sfb_reverse_fft(oo, ro, 64, 10, 10, 2500.0, store_flmn=True)
    _cmp(c, co, f, fo, "fixture vs oracle")
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.pythonrev.c_ln"))
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=2e-6)


@pytest.mark.parametrize("alias", [True, False])
def test_ring_alias_quirk_golden(handle, golden_dir, alias):
    # the numpy.fft golden of the unconjugated fold (polar rings with l > n_ring / 2)
    g = np.load(os.path.join(golden_dir, "alias_quirk.npz"))
    ns, lmax, nmax, rmax = int(g["nside"]), int(g["lmax"]), int(g["nmax"]), float(g["rmax"])
    print(f"case: ring path, alias quirk golden nside={ns} alias_compat={alias}", flush=True)
    off, rad = _pix(handle, g["points"], ns)
    _, c, _ = _hp(handle, off, rad, ns, lmax, nmax, rmax, precision=capi.PREC_MIXED, bessel=capi.BESSEL_EXACT,
                  zeros=capi.ZEROS_CONVERGED, alias_compat=alias)
    np.testing.assert_allclose(c, g["c_as_written" if alias else "c_conjugated"], rtol=TOL_LOW_L)


def test_ring_repeat_and_handle_reuse(handle):
    # deterministic run to run; the ring tables are rebuilt when nside or lmax change
    pts = workloads.uniform_ball(30000, 500.0, seed=61)
    print("case: ring path determinism and table reuse", flush=True)
    for nside, lmax in [(16, 24), (8, 36), (16, 24)]:
        off, rad = _pix(handle, pts, nside)
        _, c0, f0 = _hp(handle, off, rad, nside, lmax, 20, 500.0, precision=capi.PREC_MIXED, k_max=0.08)
        _, c1, f1 = _hp(handle, off, rad, nside, lmax, 20, 500.0, precision=capi.PREC_MIXED, k_max=0.08)
        np.testing.assert_array_equal(c0, c1)
        _, c64, f64 = _hp(handle, off, rad, nside, lmax, 20, 500.0, k_max=0.08, bessel=capi.BESSEL_EXACT)
        _cmp(c0, c64, f0, f64, f"nside {nside} lmax {lmax}")


def test_ring_rejects_other_operands(handle):
    pts = workloads.uniform_ball(2000, 100.0, seed=62)
    off, rad = _pix(handle, pts, 4)
    with pytest.raises(capi.SfbError, match="split-FP16"):
        capi.sfb_healpix_torch(handle, off, rad, 4, 8, 4, 100.0, precision=capi.PREC_MIXED,
                               mixed_operands=capi.OPERANDS_FP32)


def test_ring_low_l_in_fp64(handle):
    # mixed_fp64_below = 8: l < 8 in the FP64 pixel kernel, every C_l within 1e-6
    pts = workloads.radial_selection(60000, 800.0, r0=300.0, seed=58)
    print("case: ring path, l < 8 in FP64, nside 8 lmax 30", flush=True)
    off, rad = _pix(handle, pts, 8)
    kw = dict(k_max=0.1, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _hp(handle, off, rad, 8, 30, 40, 800.0, **kw)
    _, cr, fr = _hp(handle, off, rad, 8, 30, 40, 800.0, precision=capi.PREC_MIXED, mixed_fp64_below=8, **kw)
    assert _cmp(cr, c64, fr, f64, "l < 8 in FP64", low_l_tol=1e-9) <= TOL
