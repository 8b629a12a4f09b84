"""Batched direct SFB (sfbhip_direct_batched): several catalogues with their
own sizes, lmax, nmax, rmax, k_max and window volume in one call, each
checked against the oracle (FP64) and against the FP64 batch (mixed)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

SPECS = [  # (N, lmax, nmax, rmax, k_max, window volume, generator)
    (3000, 12, 8, 300.0, 0.0, 0.0, "uniform"),
    (20000, 40, 30, 800.0, 0.12, 0.0, "clustered"),
    (517, 9, 5, 150.0, 0.0, 2.5e7, "uniform"),
    (9000, 33, 20, 1200.0, 0.05, 0.0, "uniform"),
    (1, 6, 4, 50.0, 0.0, 0.0, "uniform"),
]


def _catalogs():
    out = []
    for i, (n, lmax, nmax, rmax, kmax, vol, gen) in enumerate(SPECS):
        if gen == "uniform":
            pts = workloads.uniform_ball(n, rmax, seed=40 + i)
        else:
            pts = workloads.clustered_ball(n // 40, rmax, mean_children=40, sigma=6.0, seed=40 + i)
        out.append(dict(np=pts, points=torch.from_numpy(pts).cuda(), lmax=lmax, nmax=nmax, rmax=rmax, k_max=kmax,
                        window_volume=vol))
    return out


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


def test_batched_fp64_vs_oracle(handle):
    cats = _catalogs()
    print("case: batched FP64, 5 catalogues", flush=True)
    res = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_FP64)
    torch.cuda.synchronize()
    for c, (k, cl, f) in zip(cats, res):
        nl = O.count_nl(c["lmax"], c["nmax"], c["rmax"], c["k_max"]) if c["k_max"] > 0 else None
        ko, co, fo = O.sfb_raw(c["np"], c["lmax"], c["nmax"], c["rmax"], nl=nl, volume=c["window_volume"],
                               store_flmn=True)
        np.testing.assert_allclose(k.cpu().numpy(), ko, rtol=1e-14)
        np.testing.assert_allclose(cl.cpu().numpy(), co, rtol=1e-10)
        assert _flmn_err(capi.split_flmn(f.cpu().numpy(), c["lmax"], c["nmax"]), fo) < 1e-10


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_batched_mixed_vs_fp64_and_single(handle, kern):
    cats = _catalogs()
    print(f"case: batched mixed vs batched FP64 and single calls, kernel {kern}", flush=True)
    r64 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_FP64)
    rmx = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_MIXED, mixed_kernel=kern)
    torch.cuda.synchronize()
    for c, (k64, c64, f64), (km, cm, fm) in zip(cats, r64, rmx):
        a, b = cm.cpu().numpy(), c64.cpu().numpy()
        assert np.all((a == 0) == (b == 0))
        m = b != 0
        assert np.max(np.abs(a[m] / b[m] - 1)) < 1e-6
        L, N = c["lmax"], c["nmax"]
        assert _flmn_err(capi.split_flmn(fm.cpu().numpy(), L, N), capi.split_flmn(f64.cpu().numpy(), L, N)) < 1e-6
        # the same catalogue alone through sfbhip_direct
        ks, cs, _ = capi.sfb_direct_torch(handle, c["points"], L, N, c["rmax"], k_max=c["k_max"],
                                          window_volume=c["window_volume"], precision=capi.PREC_MIXED,
                                          bessel=capi.BESSEL_EXACT, mixed_kernel=kern)
        s = cs.cpu().numpy()
        assert np.max(np.abs(a[m] / s[m] - 1)) < 1e-6


def test_batched_errors(handle):
    cats = _catalogs()[:2]
    with pytest.raises(ValueError, match="ncat"):
        capi.sfb_direct_batched_torch(handle, [], precision=capi.PREC_FP64)
    bad = dict(cats[1])
    bad["points"] = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="catalogue 1"):
        capi.sfb_direct_batched_torch(handle, [cats[0], bad], precision=capi.PREC_FP64)


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_batched_mixed_first_radius_not_smallest(handle, kern):
    # the batched sort orders by r with its lowest 7 bits dropped, so within a
    # key group points keep their input order.  Here the first key group holds
    # radii on both sides of a Bessel node of the l = 4, n = 0 row (t = k r / h
    # an integer, h = 1/16) and the input is in descending order: the item's
    # first point then lies above the node while the group's smallest lies
    # below.  The mixed kernel's node window must still cover the smallest.
    lmax, nmax, rmax = 8, 6, 500.0
    ball = workloads.uniform_ball(2000, rmax, seed=77)
    k, _, _ = capi.sfb_direct_torch(handle, torch.from_numpy(ball).cuda(), lmax, nmax, rmax,
                                    precision=capi.PREC_FP64)
    khr = float(k.cpu().numpy()[4, 0]) * 16.0
    for m in range(40, 400):
        r = m / khr
        while r * khr < m:
            r = np.nextafter(r, np.inf)
        while np.nextafter(r, -np.inf) * khr >= m:
            r = np.nextafter(r, -np.inf)
        b = int(np.float64(r).view(np.int64))  # smallest radius with t >= m
        if b % 128 >= 16 and r < 0.9 * rmax:
            break
    g = b - b % 128
    bits = np.arange(b + 40, g - 1, -1, dtype=np.int64)  # descending; the lowest group starts at g
    radii = bits.view(np.float64)
    assert radii[-1] * khr < m <= radii[0] * khr
    rng = np.random.default_rng(5)
    pts = np.stack([radii, rng.uniform(0.1, 3.0, radii.size), rng.uniform(0.0, 6.2, radii.size)], axis=1)
    print(f"case: batched mixed, {radii.size} radii around node {m} of l=4 n=0", flush=True)
    cats = [dict(points=torch.from_numpy(pts).cuda(), lmax=lmax, nmax=nmax, rmax=rmax, k_max=0.0,
                 window_volume=0.0),
            dict(points=torch.from_numpy(ball).cuda(), lmax=lmax, nmax=nmax, rmax=rmax, k_max=0.0,
                 window_volume=0.0)]
    r64 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_FP64)
    rmx = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_MIXED, mixed_kernel=kern)
    torch.cuda.synchronize()
    for (_, _, f64), (_, _, fm) in zip(r64, rmx):
        a = capi.split_flmn(fm.cpu().numpy(), lmax, nmax)
        e = capi.split_flmn(f64.cpu().numpy(), lmax, nmax)
        assert _flmn_err(a[2:], e[2:]) < 1e-6
