"""The benchmark configurations at full size on the GPU.

* clustered 1M points, lmax 64, k_max 0.2, rmax 1000 (993,142 points);
* HEALPix: 4,000,000 points with p(r) ~ r^2 exp(-(r/400)^2), nside 256, lmax 128, k_max 0.15;
* 64 batched catalogues (N log-uniform in [1e4, 2e6], rmax U[300, 1500], lmax U{32..128}, k_max 0.15);
* lognormal 2^25 points, lmax 256, k_max 0.5, rmax 1000 (n_l 159 -> 52).

Each uses the same seeded generator as bench.py.  The mixed path is held to the
stated tolerance against the plain FP64 GPU path on the same points (max
relative C_l error <= 1e-6, per-l relative f_lmn norm error <= 1e-6), and the
FP64 GPU path is tied to the oracle (sfb_raw.cpp:61-96, sfb_reverse_fft.cpp:127-153)
on seeded subsets of the same catalogue at the same lmax / k_max, so every
(l, n) of the configuration -- l up to 255, n up to 158 -- is checked.
"""
import numpy as np
import pytest
import torch

import bench
from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-6      # stated tolerance of the mixed path against FP64
TOL_FP64 = 1e-10      # FP64 GPU vs oracle: same algorithm, summation order only


def _cmp_mixed(cm, c64, fm, f64, lmax, nmax, what):
    a, b = cm, c64
    assert np.all((a == 0) == (b == 0)), f"{what}: zero pattern of c_ln differs"
    mask = b != 0
    rel = np.abs(a[mask] / b[mask] - 1)
    print(f"{what}: max rel C_l {rel.max():.3e}, median {np.median(rel):.3e}", flush=True)
    assert rel.max() <= TOL_MIXED
    if fm is not None:
        fa, fb = capi.split_flmn(fm, lmax, nmax), capi.split_flmn(f64, lmax, nmax)
        ferr = [np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(fa, fb) if np.linalg.norm(y) > 0]
        print(f"{what}: max per-l rel f_lmn norm {max(ferr):.3e}", flush=True)
        assert max(ferr) <= TOL_MIXED


def _direct(handle, d, lmax, nmax, rmax, store=True, **kw):
    k, c, f = capi.sfb_direct_torch(handle, d, lmax, nmax, rmax, store_flmn=store, **kw)
    torch.cuda.synchronize()
    return k.cpu().numpy(), c.cpu().numpy(), (f.cpu().numpy() if f is not None else None)


def _subset(pts, n, seed):
    rng = np.random.default_rng(seed)
    return pts[np.sort(rng.choice(pts.shape[0], n, replace=False))]


def _fp64_vs_oracle(handle, pts, cfg, nl, n, seed, what):
    sub = _subset(pts, n, seed)
    d = torch.from_numpy(sub).cuda()
    L, NM, R = cfg["lmax"], cfg["nmax"], cfg["rmax"]
    k, c, f = _direct(handle, d, L, NM, R, k_max=cfg["k_max"], precision=capi.PREC_FP64,
                      bessel=capi.BESSEL_EXACT)
    ko, co, fo = O.sfb_raw(sub, L, NM, R, nl=nl, store_flmn=True, nthreads=16)
    np.testing.assert_allclose(k, ko, rtol=1e-14)
    mask = co != 0
    assert np.all((c == 0) == ~mask)
    rel = np.abs(c[mask] / co[mask] - 1)
    print(f"{what}: FP64 GPU vs oracle on {n} points: max rel C_l {rel.max():.3e}", flush=True)
    assert rel.max() < TOL_FP64
    fa = capi.split_flmn(f, L, NM)
    err = max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(fa, fo) if np.abs(y).max() > 0)
    assert err < TOL_FP64


@pytest.mark.timeout(600)
def test_clustered1m(handle):
    cfg = bench.WORKLOADS["clustered1m"]
    print("case: clustered1m, 993,142 points, lmax 64, k_max 0.2", flush=True)
    pts = bench.make_points("clustered1m")
    L, NM, R = cfg["lmax"], cfg["nmax"], cfg["rmax"]
    nl = np.array(O.count_nl(L, NM, R, cfg["k_max"]))
    d = torch.from_numpy(pts).cuda()
    _, cm, fm = _direct(handle, d, L, NM, R, k_max=cfg["k_max"], precision=capi.PREC_MIXED,
                        bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _direct(handle, d, L, NM, R, k_max=cfg["k_max"], precision=capi.PREC_FP64,
                          bessel=capi.BESSEL_EXACT)
    _cmp_mixed(cm, c64, fm, f64, L, NM, "clustered1m mixed vs FP64")
    _fp64_vs_oracle(handle, pts, cfg, nl, 16384, 7, "clustered1m")


@pytest.mark.timeout(900)
def test_lognormal2p25(handle):
    cfg = bench.WORKLOADS["lognormal2p25"]
    print("case: lognormal2p25, 2^25 points, lmax 256, k_max 0.5", flush=True)
    pts = bench.make_points("lognormal2p25")
    L, NM, R = cfg["lmax"], cfg["nmax"], cfg["rmax"]
    nl = np.array(O.count_nl(L, NM, R, cfg["k_max"]))
    assert nl[0] == 159 and nl[-1] == 52
    d = torch.from_numpy(pts).cuda()
    print("lognormal2p25: mixed call", flush=True)
    _, cm, fm = _direct(handle, d, L, NM, R, k_max=cfg["k_max"], precision=capi.PREC_MIXED,
                        bessel=capi.BESSEL_EXACT)
    print("lognormal2p25: FP64 call", flush=True)
    _, c64, f64 = _direct(handle, d, L, NM, R, k_max=cfg["k_max"], precision=capi.PREC_FP64,
                          bessel=capi.BESSEL_EXACT)
    del d
    _cmp_mixed(cm, c64, fm, f64, L, NM, "lognormal2p25 mixed vs FP64")
    # the rows of the third 64-row group (n >= 128) and the highest multipoles separately
    a, b = cm[:, 128:], c64[:, 128:]
    assert np.count_nonzero(b) > 0
    assert np.abs(a[b != 0] / b[b != 0] - 1).max() <= TOL_MIXED
    a, b = cm[200:], c64[200:]
    assert np.abs(a[b != 0] / b[b != 0] - 1).max() <= TOL_MIXED
    _fp64_vs_oracle(handle, pts, cfg, nl, 1024, 9, "lognormal2p25")


@pytest.mark.timeout(600)
def test_batched64(handle):
    print("case: batched64, 64 catalogues", flush=True)
    batch = bench.make_batch()
    cats = [dict(points=torch.from_numpy(c["np"]).cuda(), lmax=c["lmax"], nmax=c["nmax"], rmax=c["rmax"],
                 k_max=c["k_max"]) for c in batch]
    rmx = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_MIXED)
    r64 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_FP64,
                                        bessel=capi.BESSEL_EXACT)
    torch.cuda.synchronize()
    worst = 0.0
    for i, (c, (km, cm, fm), (k64, c64, f64)) in enumerate(zip(batch, rmx, r64)):
        np.testing.assert_array_equal(km.cpu().numpy(), k64.cpu().numpy())
        _cmp_mixed(cm.cpu().numpy(), c64.cpu().numpy(), fm.cpu().numpy(), f64.cpu().numpy(), c["lmax"], c["nmax"],
                   f"batched64 catalogue {i} (N={c['np'].shape[0]}, lmax {c['lmax']})")
        b = c64.cpu().numpy()
        worst = max(worst, float(np.abs(cm.cpu().numpy()[b != 0] / b[b != 0] - 1).max()))
    print(f"batched64: worst C_l over the batch {worst:.3e}", flush=True)
    # the batch equals single mixed calls of the same catalogues to the same tolerance
    for i in (0, 17, 40, 63):
        c = batch[i]
        _, cs, fs = _direct(handle, cats[i]["points"], c["lmax"], c["nmax"], c["rmax"], k_max=c["k_max"],
                            precision=capi.PREC_MIXED, bessel=capi.BESSEL_EXACT)
        _cmp_mixed(cs, r64[i][1].cpu().numpy(), fs, r64[i][2].cpu().numpy(), c["lmax"], c["nmax"],
                   f"batched64 catalogue {i} single mixed call vs FP64 batch")
        np.testing.assert_allclose(cs, rmx[i][1].cpu().numpy(), rtol=2 * TOL_MIXED)


@pytest.mark.timeout(600)
def test_healpix4m(handle):
    cfg = bench.WORKLOADS["healpix4m"]
    L, NM, R, ns = cfg["lmax"], cfg["nmax"], cfg["rmax"], cfg["nside"]
    print("case: healpix4m, 4,000,000 points, nside 256, lmax 128, k_max 0.15", flush=True)
    pts = workloads.radial_selection(4_000_000, R, r0=400.0, seed=3)
    off, rad = capi.pixelize_torch(handle, torch.from_numpy(pts).cuda(), ns)
    torch.cuda.synchronize()
    oo, ro = O.pixelize(pts, ns)
    np.testing.assert_array_equal(off.cpu().numpy(), oo)
    np.testing.assert_array_equal(rad.cpu().numpy(), ro)
    print("healpix4m: binning bit-exact", flush=True)
    k, c, f = capi.sfb_healpix_torch(handle, off, rad, ns, L, NM, R, k_max=cfg["k_max"], store_flmn=True)
    torch.cuda.synchronize()
    c = c.cpu().numpy()
    nl = np.array(O.count_nl(L, NM, R, cfg["k_max"]))
    sample = list(range(8, L, 16))
    nls = np.zeros_like(nl)
    nls[sample] = nl[sample]
    ko, co, fo = O.sfb_reverse_fft(oo, ro, ns, L, NM, R, nl=nls, interpolated=True, nthreads=16, store_flmn=True)
    fa = capi.split_flmn(f.cpu().numpy(), L, NM)
    for l in sample:
        m = co[l] != 0
        assert m.sum() == nl[l]
        np.testing.assert_allclose(c[l][m], co[l][m], rtol=TOL_FP64)
        np.testing.assert_allclose(fa[l][:, :nl[l]], fo[l][:, :nl[l]], rtol=0,
                                   atol=TOL_FP64 * np.abs(fo[l]).max())
    print(f"healpix4m: FP64 HEALPix path vs oracle at l = {sample}", flush=True)


@pytest.mark.timeout(600)
def test_healpix4m_ring_path(handle):
    # the ring path (FP32 maps, ring Legendre table, split-FP16 ring contraction)
    # at full size against the FP64 HEALPix path with exact j_l, every (l, n)
    cfg = bench.WORKLOADS["healpix4m"]
    L, NM, R, ns = cfg["lmax"], cfg["nmax"], cfg["rmax"], cfg["nside"]
    print("case: healpix4m ring path vs FP64 exact, nside 256, lmax 128", flush=True)
    pts = workloads.radial_selection(4_000_000, R, r0=400.0, seed=3)
    off, rad = capi.pixelize_torch(handle, torch.from_numpy(pts).cuda(), ns)
    kw = dict(k_max=cfg["k_max"], store_flmn=True)
    _, cr, fr = capi.sfb_healpix_torch(handle, off, rad, ns, L, NM, R, precision=capi.PREC_MIXED, **kw)
    _, ce, fe = capi.sfb_healpix_torch(handle, off, rad, ns, L, NM, R, precision=capi.PREC_FP64,
                                       bessel=capi.BESSEL_EXACT, **kw)
    torch.cuda.synchronize()
    # the ring path's stated tolerance: C_l <= 1e-6 for l >= 8, <= 5e-6 for 2 <= l < 8
    # (l < 2 run in FP64), per-l f_lmn <= 1e-6
    a, b = cr.cpu().numpy(), ce.cpu().numpy()
    assert np.all((a == 0) == (b == 0))
    m = b != 0
    full = np.where(m, np.abs(a / np.where(m, b, 1) - 1), 0)
    print(f"healpix4m ring vs FP64 exact: max rel C_l l < 8 {full[:8].max():.3e}, l >= 8 {full[8:].max():.3e}, "
          f"median {np.median(full[m]):.3e}", flush=True)
    assert full[8:].max() <= TOL_MIXED and full[:8].max() <= 5 * TOL_MIXED
    fa, fb = capi.split_flmn(fr.cpu().numpy(), L, NM), capi.split_flmn(fe.cpu().numpy(), L, NM)
    ferr = [np.linalg.norm(x - y) / np.linalg.norm(y) for x, y in zip(fa, fb) if np.linalg.norm(y) > 0]
    print(f"healpix4m ring: max per-l rel f_lmn norm {max(ferr):.3e}", flush=True)
    assert max(ferr) <= TOL_MIXED
