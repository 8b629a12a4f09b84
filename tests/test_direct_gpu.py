"""Parity of the FP64 direct path (libcatana_hip.so through the C-ABI) with
the CPU oracle and the reference's golden fixtures."""
import os

import numpy as np
import torch  # noqa: F401  (imported at collection: the first import can take minutes)
import pytest

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

CF = 5535184878.03 / (2500 ** 3 * 4 / 3 * np.pi)  # test/python/decomposition_test.py:15


def _run(handle, pts, lmax, nmax, rmax, store=True, **kw):
    import torch
    d = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    k, c, f = capi.sfb_direct_torch(handle, d, lmax, nmax, rmax, store_flmn=store, **kw)
    torch.cuda.synchronize()
    fl = capi.split_flmn(f.cpu().numpy(), lmax, nmax) if store else None
    return k.cpu().numpy(), c.cpu().numpy(), fl


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


def test_gaussian_fixture_exact(handle, gaussian_points, golden_dir):
    print("case: gaussian fixture, lmax=nmax=10, rmax=2500, exact bessel", flush=True)
    k, c, f = _run(handle, gaussian_points, 10, 10, 2500.0, bessel=capi.BESSEL_EXACT)
    ko, co, fo = O.sfb_raw(gaussian_points, 10, 10, 2500.0, store_flmn=True)
    np.testing.assert_allclose(k, ko, rtol=1e-13)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10
    ref = np.loadtxt(os.path.join(golden_dir, "gaussian_catalog.python.c_ln"))
    np.testing.assert_allclose(c * CF ** 2, ref, rtol=2e-6)


def test_uniform_ball_exact_vs_oracle(handle):
    pts = workloads.uniform_ball(3000, 500.0, seed=3)
    print("case: uniform ball N=3000 lmax=20 nmax=12 exact", flush=True)
    k, c, f = _run(handle, pts, 20, 12, 500.0, bessel=capi.BESSEL_EXACT)
    ko, co, fo = O.sfb_raw(pts, 20, 12, 500.0, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10


def test_uniform_ball_reference_table(handle):
    pts = workloads.uniform_ball(3000, 500.0, seed=4)
    print("case: uniform ball N=3000 lmax=12 nmax=8 reference linear table", flush=True)
    k, c, f = _run(handle, pts, 12, 8, 500.0, bessel=capi.BESSEL_AUTO)
    ko, co, fo = O.sfb_raw(pts, 12, 8, 500.0, interpolated=True, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10


def test_kmax_cut(handle):
    pts = workloads.uniform_ball(2000, 500.0, seed=5)
    print("case: k_max=0.1 rmax=500 lmax=16 nmax=20", flush=True)
    k, c, f = _run(handle, pts, 16, 20, 500.0, k_max=0.1, bessel=capi.BESSEL_EXACT)
    nl = O.count_nl(16, 20, 500.0, 0.1)
    ko, co, fo = O.sfb_raw(pts, 16, 20, 500.0, nl=nl, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10, atol=0)
    for l in range(16):
        assert np.all(c[l, nl[l]:] == 0)
    assert nl[0] == 15 and nl[15] == 9
