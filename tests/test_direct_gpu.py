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


def _rel(a, b):
    m = b != 0
    return np.abs(a[m] / b[m] - 1)


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_mixed_vs_fp64_uniform(handle, kern):
    pts = workloads.uniform_ball(20000, 500.0, seed=21)
    print(f"kernel {kern} case: mixed vs fp64, uniform N=20000 lmax=24 k_max=0.15 rmax=500", flush=True)
    kw = dict(k_max=0.15, bessel=capi.BESSEL_EXACT)
    k64, c64, f64 = _run(handle, pts, 24, 30, 500.0, precision=capi.PREC_FP64, **kw)
    km, cm, fm = _run(handle, pts, 24, 30, 500.0, precision=capi.PREC_MIXED, mixed_kernel=kern, **kw)
    np.testing.assert_array_equal(km, k64)
    rel = _rel(cm, c64)
    print("max rel c_ln", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-6
    assert _flmn_err(fm, f64) < 1e-6
    assert np.all((cm == 0) == (c64 == 0))


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_mixed_all_l_in_mixed_kernel(handle, kern):
    # mixed_fp64_below = 0: every l through the FP32-operand kernel; f_lmn stays
    # within 1e-6 of FP64 relative to each l block's largest coefficient
    pts = workloads.uniform_ball(20000, 500.0, seed=24)
    print(f"kernel {kern} case: mixed kernel for all l, uniform N=20000 lmax=24", flush=True)
    kw = dict(k_max=0.15, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _run(handle, pts, 24, 30, 500.0, precision=capi.PREC_FP64, **kw)
    _, cm, fm = _run(handle, pts, 24, 30, 500.0, precision=capi.PREC_MIXED, mixed_kernel=kern, mixed_fp64_below=0, **kw)
    assert _flmn_err(fm, f64) < 1e-6
    assert np.median(_rel(cm, c64)) < 2e-7


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_mixed_vs_oracle_points_beyond_rmax(handle, kern):
    # r up to 1.3 rmax: Bessel arguments past the table fall back to FP64 j_l
    pts = workloads.uniform_ball(4000, 650.0, seed=22)
    print(f"kernel {kern} case: mixed, points beyond rmax, N=4000 lmax=10 nmax=8 rmax=500", flush=True)
    k, c, f = _run(handle, pts, 10, 8, 500.0, precision=capi.PREC_MIXED, mixed_kernel=kern, bessel=capi.BESSEL_EXACT)
    ko, co, fo = O.sfb_raw(pts, 10, 8, 500.0, store_flmn=True)
    assert _rel(c, co).max() < 2e-6
    assert _flmn_err(f, fo) < 2e-6


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_mixed_vs_fp64_clustered_high_l(handle, kern):
    pts = workloads.clustered_ball(1500, 1000.0, mean_children=40, sigma=5.0, seed=23)
    print(f"kernel {kern} case: mixed vs fp64, clustered N={pts.shape[0]} lmax=140 k_max=0.2 rmax=1000", flush=True)
    kw = dict(k_max=0.2, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _run(handle, pts, 140, 64, 1000.0, precision=capi.PREC_FP64, **kw)
    _, cm, fm = _run(handle, pts, 140, 64, 1000.0, precision=capi.PREC_MIXED, mixed_kernel=kern, **kw)
    rel = _rel(cm, c64)
    print("max rel c_ln", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-6
    assert _flmn_err(fm, f64) < 1e-6


@pytest.mark.parametrize("kern", [capi.MIXED_KERNEL_256, capi.MIXED_KERNEL_512, capi.MIXED_KERNEL_GROUPED])
def test_mixed_vs_fp64_two_row_groups(handle, kern):
    # n_l up to 76 > 64: multipoles with a second 64-row group (its 12 rows run
    # with the tile's points split four ways) and, for l >= 64, items whose m
    # chunks start below m = l
    pts = workloads.uniform_ball(3000, 600.0, seed=26)
    print(f"kernel {kern} case: mixed vs fp64, uniform N=3000 lmax=80 k_max=0.4 rmax=600 (n_l up to 76)", flush=True)
    kw = dict(k_max=0.4, bessel=capi.BESSEL_EXACT)
    k64, c64, f64 = _run(handle, pts, 80, 80, 600.0, precision=capi.PREC_FP64, **kw)
    km, cm, fm = _run(handle, pts, 80, 80, 600.0, precision=capi.PREC_MIXED, mixed_kernel=kern, **kw)
    assert np.count_nonzero(c64[:, 64]) > 0  # some l has a 65th mode
    np.testing.assert_array_equal(km, k64)
    rel = _rel(cm, c64)
    print("max rel c_ln", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-6
    assert _flmn_err(fm, f64) < 1e-6
    assert np.all((cm == 0) == (c64 == 0))


def test_mixed_fp64_multipoles_tables(handle):
    # mixed_fp64_below = lmax: every l in FP64 MFMA kernels with the quintic j_l
    # table (error ~1e-10 of the amplitude) -- l < 8 in the low-l kernel,
    # 8 <= l < 12 in the general FP64 kernel; points up to 1.2 rmax
    pts = workloads.clustered_ball(400, 600.0, mean_children=30, sigma=8.0, seed=25)
    print(f"case: FP64 multipoles with j_l tables, N={pts.shape[0]} lmax=12 k_max=0.3 rmax=500", flush=True)
    kw = dict(k_max=0.3, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _run(handle, pts, 12, 60, 500.0, precision=capi.PREC_FP64, **kw)
    _, cq, fq = _run(handle, pts, 12, 60, 500.0, precision=capi.PREC_MIXED, mixed_fp64_below=12, **kw)
    rlo, rhi = _rel(cq[:8], c64[:8]).max(), _rel(cq[8:], c64[8:]).max()
    lo, hi = _flmn_err(fq[:8], f64[:8]), _flmn_err(fq[8:], f64[8:])
    print("max rel c_ln l<8", rlo, "l>=8", rhi, "f_lmn", lo, hi)
    assert rlo < 1e-7 and lo < 1e-8
    assert rhi < 1e-7 and hi < 1e-8


@pytest.mark.parametrize("ops,chunks,kern", [(capi.OPERANDS_FP32, 4, 0), (capi.OPERANDS_SPLIT_FP16, 1, 1),
                                             (capi.OPERANDS_SPLIT_FP16, 8, 1), (capi.OPERANDS_SPLIT3_FP16, 2, 0),
                                             (capi.OPERANDS_SPLIT_FP16, 1, 2), (capi.OPERANDS_SPLIT_FP16, 2, 2),
                                             (capi.OPERANDS_SPLIT_FP16, 4, 2), (capi.OPERANDS_SPLIT_FP16, 1, 3),
                                             (capi.OPERANDS_SPLIT_FP16, 2, 3), (capi.OPERANDS_SPLIT_FP16, 8, 3)])
def test_mixed_operand_formats_and_chunks(handle, ops, chunks, kern):
    # every operand format and work-item chunk count against FP64; l up to 47
    # spans 3 chunks, so chunks=1 and 2 split the m range over several passes
    pts = workloads.clustered_ball(600, 800.0, mean_children=40, sigma=6.0, seed=26)
    print(f"case: operands={ops} chunks={chunks} kernel={kern} N={pts.shape[0]} lmax=48 k_max=0.2 rmax=800",
          flush=True)
    kw = dict(k_max=0.2, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _run(handle, pts, 48, 60, 800.0, precision=capi.PREC_FP64, **kw)
    _, cm, fm = _run(handle, pts, 48, 60, 800.0, precision=capi.PREC_MIXED, mixed_operands=ops,
                     mixed_chunks=chunks, mixed_kernel=kern, **kw)
    rel = _rel(cm, c64)
    print("max rel c_ln", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-6
    assert _flmn_err(fm, f64) < 1e-6


def test_device_jl_vs_mpmath_goldens(handle, golden_dir):
    # one-point catalogues place k_{l,0} r at each golden argument x of special_functions.npz
    # (mpmath j_l, 450 digits): c_l0 = (2/pi) V^2 k^2 j_l(x)^2 / (4 pi) gives the device's |j_l(x)|
    # directly, not through the oracle's j_l
    g = np.load(os.path.join(golden_dir, "special_functions.npz"))
    ls, xs, vs = g["jl_l"], g["jl_x"], g["jl_v"]
    print(f"case: device j_l at {ls.size} golden arguments", flush=True)
    cats, meta = [], []
    for l, x, v in zip(ls, xs, vs):
        z0 = O.bessel_zeros(int(l), 1)[0]
        rmax = 1.0
        r = x / z0
        cats.append(dict(points=torch.tensor([[r, 0.9, 1.3]], dtype=torch.float64, device="cuda"),
                         lmax=int(l) + 1, nmax=1, rmax=rmax))
        meta.append((int(l), x, v, rmax))
    got = []
    for i in range(0, len(cats), 128):
        res = capi.sfb_direct_batched_torch(handle, cats[i:i + 128], precision=capi.PREC_FP64)
        torch.cuda.synchronize()
        got += [(k.cpu().numpy(), c.cpu().numpy()) for k, c, _ in res]
    checked = 0
    for (l, x, v, rmax), (k, c) in zip(meta, got):
        V = 4.0 / 3.0 * np.pi * rmax ** 3
        kk = k[l, 0]
        j = np.sqrt(c[l, 0] * 4 * np.pi / (2 / np.pi * V ** 2 * kk ** 2))
        if abs(v) > 1e-250:
            # r = x / z0 moves x by about an ulp: |j_l'| <= 1 bounds the change by ~x 2^-52
            np.testing.assert_allclose(j, abs(v), rtol=1e-12, atol=4e-16 * max(x, 1.0))
            checked += 1
        else:
            assert j <= 1e-240
    assert checked > 250


@pytest.mark.parametrize("n,kern", [(400, capi.MIXED_KERNEL_512), (400, capi.MIXED_KERNEL_256),
                                    (6000, capi.MIXED_KERNEL_512), (400, capi.MIXED_KERNEL_GROUPED),
                                    (6000, capi.MIXED_KERNEL_GROUPED)])
def test_mixed_lmax256_third_row_group_and_global_table(handle, n, kern):
    # lmax 256, k_max 0.5, rmax 1000: n_l from 159 down to 52, so l >= 141 and rows
    # n >= 128 (a third 64-row group, or the rows after a 128-row item) are
    # exercised; with N = 400 one work item spans radii 0 .. rmax, its Bessel
    # arguments cover more intervals than LDS holds and it reads the global table
    pts = workloads.uniform_ball(n, 1000.0, seed=27)
    print(f"kernel {kern} case: mixed vs fp64, uniform N={n} lmax=256 k_max=0.5 rmax=1000", flush=True)
    kw = dict(k_max=0.5, bessel=capi.BESSEL_EXACT)
    _, c64, f64 = _run(handle, pts, 256, 160, 1000.0, precision=capi.PREC_FP64, **kw)
    _, cm, fm = _run(handle, pts, 256, 160, 1000.0, precision=capi.PREC_MIXED, mixed_kernel=kern, **kw)
    assert np.count_nonzero(c64[:, 128:]) > 0 and np.count_nonzero(c64[141:]) > 0
    rel = _rel(cm, c64)
    print("max rel c_ln", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-6
    assert _rel(cm[:, 128:], c64[:, 128:]).max() < 1e-6
    assert _flmn_err(fm, f64) < 1e-6
    assert np.all((cm == 0) == (c64 == 0))


@pytest.mark.parametrize("prec", [capi.PREC_MIXED, capi.PREC_FP64])
def test_plan_kept_between_calls(handle, prec):
    # sfbhip_direct keeps its last plan (work items on the device) for calls of the
    # same shape: the second catalogue of the same size and radius bound reuses it,
    # other sizes and multipole counts in between replace it; every result must be
    # bit-identical to what a fresh handle computes for the same call
    a = workloads.uniform_ball(5000, 500.0, seed=21)
    b = workloads.uniform_ball(5000, 500.0, seed=22)
    c = workloads.uniform_ball(4000, 500.0, seed=23)
    seq = [(a, 40), (b, 40), (c, 40), (b, 40), (a, 24), (b, 40)]
    print(f"case: plan reuse, precision {prec}, N = 5000/5000/4000, lmax 40/24, nmax 16, rmax 500", flush=True)
    got = [_run(handle, p, L, 16, 500.0, precision=prec) for p, L in seq]
    fresh = capi.Handle(0)
    try:
        for (p, L), g in zip(seq, got):
            fresh.close()
            fresh = capi.Handle(0)
            ref = _run(fresh, p, L, 16, 500.0, precision=prec)
            np.testing.assert_array_equal(g[0], ref[0])
            np.testing.assert_array_equal(g[1], ref[1])
            for x, y in zip(g[2], ref[2]):
                np.testing.assert_array_equal(x, y)
    finally:
        fresh.close()
