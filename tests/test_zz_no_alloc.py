"""sfbhip_direct_noalloc: the FP64 direct decomposition with every device
buffer carved from the library's static per-GPU array.  Checked through the
C-ABI against sfbhip_direct (FP64, exact j_l) on two calls in a row, the
second with a smaller catalogue: same plan, same kernels and a fixed-order
reduction, so the results are bit-identical (tolerance 0)."""
import ctypes

import numpy as np
import pytest
import torch

from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu


def _noalloc(points, lmax, nmax, rmax, store_flmn, **kw):
    p = capi.make_params(lmax, nmax, rmax, precision=capi.PREC_FP64, bessel=capi.BESSEL_EXACT, **kw)
    k = torch.full((lmax, nmax), float("nan"), dtype=torch.float64, device=points.device)
    c = torch.full((lmax, nmax), float("nan"), dtype=torch.float64, device=points.device)
    f = (torch.full((2 * capi.flmn_count(lmax, nmax),), float("nan"), dtype=torch.float64, device=points.device)
         if store_flmn else None)
    torch.cuda.synchronize()
    rc = capi.lib().sfbhip_direct_noalloc(ctypes.c_void_p(points.data_ptr()), ctypes.c_int64(points.shape[0]),
                                          ctypes.byref(p), ctypes.c_void_p(k.data_ptr()),
                                          ctypes.c_void_p(c.data_ptr()),
                                          ctypes.c_void_p(f.data_ptr()) if f is not None else None, None)
    capi.check(rc, "sfbhip_direct_noalloc")
    return k, c, f


CASES = [  # (points, lmax, nmax, rmax, k_max, store f_lmn): the second call is smaller
    ("uniform", 30000, 24, 16, 500.0, 0.05, True),
    ("clustered", 4000, 12, 10, 300.0, 0.0, False),
]


def test_noalloc_matches_direct_on_two_calls(handle):
    for gen, n, lmax, nmax, rmax, kmax, store in CASES:
        if gen == "uniform":
            pts = workloads.uniform_ball(n, rmax, seed=301)
        else:
            pts = workloads.clustered_ball(n // 40, rmax, mean_children=40, sigma=5.0, seed=302)
        P = torch.from_numpy(pts).cuda()
        print(f"case: noalloc vs direct, {gen} N={P.shape[0]} lmax={lmax} nmax={nmax} store={store}", flush=True)
        ref = capi.sfb_direct_torch(handle, P, lmax, nmax, rmax, store_flmn=store, precision=capi.PREC_FP64,
                                    bessel=capi.BESSEL_EXACT, k_max=kmax)
        got = _noalloc(P, lmax, nmax, rmax, store, k_max=kmax)
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            if b is not None:
                assert torch.equal(a, b)
        if kmax > 0:  # the k_max cut leaves zero columns
            assert (got[1] == 0).any()


def test_noalloc_refusals(handle):
    P = torch.from_numpy(workloads.uniform_ball(2000, 100.0, seed=303)).cuda()
    p = capi.make_params(8, 6, 100.0, precision=capi.PREC_MIXED)
    k = torch.empty((8, 6), dtype=torch.float64, device="cuda")
    c = torch.empty_like(k)
    L = capi.lib()
    args = (ctypes.c_void_p(P.data_ptr()), ctypes.c_int64(P.shape[0]))
    rc = L.sfbhip_direct_noalloc(*args, ctypes.byref(p), ctypes.c_void_p(k.data_ptr()),
                                 ctypes.c_void_p(c.data_ptr()), None, None)
    assert rc == capi.NOT_SUPPORTED and b"FP64" in L.sfbhip_last_error()
    p = capi.make_params(8, 6, 100.0, precision=capi.PREC_FP64)  # bessel AUTO with 2000 points: the table
    rc = L.sfbhip_direct_noalloc(*args, ctypes.byref(p), ctypes.c_void_p(k.data_ptr()),
                                 ctypes.c_void_p(c.data_ptr()), None, None)
    assert rc == capi.NOT_SUPPORTED and b"exact" in L.sfbhip_last_error()
    # more points than the static array holds: refused before any device work
    p = capi.make_params(8, 6, 100.0, precision=capi.PREC_FP64, bessel=capi.BESSEL_EXACT)
    rc = L.sfbhip_direct_noalloc(ctypes.c_void_p(P.data_ptr()), ctypes.c_int64(1 << 23), ctypes.byref(p),
                                 ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(c.data_ptr()), None, None)
    assert rc == capi.NOT_SUPPORTED and b"bytes of device workspace" in L.sfbhip_last_error()
    bad = P.clone()
    bad[5, 0] = -1.0
    with pytest.raises(ValueError, match="r < 0"):
        _noalloc(bad, 8, 6, 100.0, False)
    # a good call after the refusals still matches sfbhip_direct
    ref = capi.sfb_direct_torch(handle, P, 8, 6, 100.0, precision=capi.PREC_FP64, bessel=capi.BESSEL_EXACT)
    got = _noalloc(P, 8, 6, 100.0, False)
    torch.cuda.synchronize()
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
