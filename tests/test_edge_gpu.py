"""Edge cases of the direct path on the GPU: single point (closed form),
ragged sizes, points on the poles / origin / phi wrap, tiny lmax/nmax, l with
no modes under the k_max cut, and the error paths."""
import numpy as np
import pytest
import torch  # noqa: F401

from oracle import oracle as O
from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu


def _run(handle, pts, lmax, nmax, rmax, store=True, **kw):
    d = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float64)).cuda()
    k, c, f = capi.sfb_direct_torch(handle, d, lmax, nmax, rmax, store_flmn=store, **kw)
    torch.cuda.synchronize()
    return k.cpu().numpy(), c.cpu().numpy(), (capi.split_flmn(f.cpu().numpy(), lmax, nmax) if store else None)


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


@pytest.mark.parametrize("precision,tol", [(capi.PREC_FP64, 1e-11), (capi.PREC_MIXED, 1e-6)])
def test_single_point_closed_form(handle, precision, tol):
    # N = 1: c_ln = (2/pi) V^2 k^2 j_l(k r)^2 / (4 pi), whatever the angles (addition theorem)
    r, rmax, lmax, nmax = 123.4, 500.0, 20, 12
    pts = np.array([[r, 0.7, 2.1]])
    print(f"case: single point, precision {precision}", flush=True)
    k, c, _ = _run(handle, pts, lmax, nmax, rmax, precision=precision, bessel=capi.BESSEL_EXACT,
                   mixed_fp64_below=0)
    V = 4.0 / 3.0 * np.pi * rmax ** 3
    jl = np.array([[O.sbessel_jl(l, k[l, n] * r) for n in range(nmax)] for l in range(lmax)])
    expect = 2 / np.pi * V ** 2 * k ** 2 * jl ** 2 / (4 * np.pi)
    big = expect > (1e-12 if precision == capi.PREC_FP64 else 1e-4) * expect.max()
    np.testing.assert_allclose(c[big], expect[big], rtol=tol)


@pytest.mark.parametrize("n", [1, 2, 255, 257, 1023, 4097])
def test_ragged_sizes(handle, n):
    pts = workloads.uniform_ball(n, 300.0, seed=100 + n)
    print(f"case: ragged N={n}", flush=True)
    _, c, f = _run(handle, pts, 18, 9, 300.0, bessel=capi.BESSEL_EXACT)
    _, co, fo = O.sfb_raw(pts, 18, 9, 300.0, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-10)
    assert _flmn_err(f, fo) < 1e-10
    _, cm, fm = _run(handle, pts, 18, 9, 300.0, precision=capi.PREC_MIXED, bessel=capi.BESSEL_EXACT)
    assert _flmn_err(fm, fo) < 1e-6


def test_special_positions(handle):
    # origin, both poles, phi = 0 and phi just below 2 pi, theta near the sin(theta) clamp
    pts = workloads.uniform_ball(3000, 200.0, seed=7)
    special = np.array([[0.0, 1.0, 1.0], [50.0, 0.0, 0.0], [60.0, np.pi, 3.0], [70.0, 1.0, 0.0],
                        [80.0, 2.0, 2 * np.pi - 1e-15], [90.0, 1e-30, 1.0], [100.0, np.pi - 1e-30, 5.0]])
    pts = np.vstack([pts, special])
    print("case: special positions", flush=True)
    _, c, f = _run(handle, pts, 40, 10, 200.0, bessel=capi.BESSEL_EXACT)
    _, co, fo = O.sfb_raw(pts, 40, 10, 200.0, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-9)
    assert _flmn_err(f, fo) < 1e-9
    _, cm, fm = _run(handle, pts, 40, 10, 200.0, precision=capi.PREC_MIXED, bessel=capi.BESSEL_EXACT)
    assert _flmn_err(fm, fo) < 1e-6


def test_tiny_and_empty_modes(handle):
    pts = workloads.uniform_ball(2000, 100.0, seed=3)
    print("case: lmax=1 nmax=1; k_max leaving high l without modes", flush=True)
    _, c, f = _run(handle, pts, 1, 1, 100.0, bessel=capi.BESSEL_EXACT)
    _, co, fo = O.sfb_raw(pts, 1, 1, 100.0, store_flmn=True)
    np.testing.assert_allclose(c, co, rtol=1e-12)
    # k_max = 0.12, rmax 100: z < 12 keeps modes only for small l; l >= 11 have none
    nl = O.count_nl(30, 6, 100.0, 0.12)
    assert nl[-1] == 0 and nl[0] > 0
    for prec in (capi.PREC_FP64, capi.PREC_MIXED):
        _, c, _ = _run(handle, pts, 30, 6, 100.0, k_max=0.12, precision=prec, bessel=capi.BESSEL_EXACT)
        _, co, _ = O.sfb_raw(pts, 30, 6, 100.0, nl=nl, store_flmn=True)
        assert np.all(c[co == 0] == 0)
        np.testing.assert_allclose(c[co != 0], co[co != 0], rtol=1e-6 if prec else 1e-10)


def test_error_paths(handle):
    d = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="empty"):
        capi.sfb_direct_torch(handle, d, 4, 4, 10.0)
    bad = torch.tensor([[-1.0, 0.5, 0.5], [1.0, 0.5, 0.5]], dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="r < 0"):
        capi.sfb_direct_torch(handle, bad, 4, 4, 10.0, bessel=capi.BESSEL_EXACT)
    ok = torch.tensor([[1.0, 0.5, 0.5]], dtype=torch.float64, device="cuda")
    with pytest.raises(capi.SfbError, match="lmax <= 512"):
        capi.sfb_direct_torch(handle, ok, 600, 2, 10.0, precision=capi.PREC_MIXED)
    with pytest.raises(ValueError, match="mixed_chunks"):
        capi.sfb_direct_torch(handle, ok, 4, 2, 10.0, precision=capi.PREC_MIXED, mixed_chunks=3)
    # the handle stays usable after errors
    _, c, _ = capi.sfb_direct_torch(handle, ok, 4, 2, 10.0, bessel=capi.BESSEL_EXACT)
    assert torch.isfinite(c).all()


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
def test_negative_zero_radius(handle, precision):
    # r = -0.0 is a valid radius (it compares equal to 0): it sorts with +0.0,
    # and the catalogue gives the coefficients of the same catalogue with +0.0
    pts = workloads.uniform_ball(3000, 400.0, seed=8)
    pts[::500, 0] = 0.0
    neg = pts.copy()
    neg[::500, 0] = -0.0
    print(f"case: r = -0.0, precision {precision}", flush=True)
    _, c0, f0 = _run(handle, pts, 20, 10, 400.0, precision=precision)
    _, c1, f1 = _run(handle, neg, 20, 10, 400.0, precision=precision)
    np.testing.assert_allclose(c1, c0, rtol=1e-13, atol=0)
    assert _flmn_err(f1, f0) < 1e-13


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
def test_deterministic_run_to_run(handle, precision):
    # partial blocks are summed in a fixed order (no atomics): repeated calls,
    # and a second handle, give bit-identical k_ln, c_ln and f_lmn
    pts = workloads.clustered_ball(300, 700.0, mean_children=30, sigma=6.0, seed=31)
    print(f"case: determinism, N={pts.shape[0]}, precision {precision}", flush=True)
    kw = dict(k_max=0.25, precision=precision)
    k0, c0, f0 = _run(handle, pts, 70, 70, 700.0, **kw)
    k1, c1, f1 = _run(handle, pts, 70, 70, 700.0, **kw)
    h2 = capi.Handle(0)
    try:
        k2, c2, f2 = _run(h2, pts, 70, 70, 700.0, **kw)
    finally:
        h2.close()
    for k, c, f in ((k1, c1, f1), (k2, c2, f2)):
        np.testing.assert_array_equal(k, k0)
        np.testing.assert_array_equal(c, c0)
        for a, b in zip(f, f0):
            np.testing.assert_array_equal(a, b)


def test_batched_deterministic(handle):
    cats = []
    for i, (n, L, R) in enumerate([(4000, 30, 400.0), (900, 50, 900.0)]):
        p = workloads.uniform_ball(n, R, seed=60 + i)
        cats.append(dict(points=torch.from_numpy(p).cuda(), lmax=L, nmax=40, rmax=R, k_max=0.12))
    print("case: batched determinism", flush=True)
    r0 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_MIXED)
    r1 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_MIXED)
    torch.cuda.synchronize()
    for a, b in zip(r0, r1):
        for x, y in zip(a, b):
            assert torch.equal(x, y)


@pytest.mark.parametrize("bad_r", [-1.0, -np.inf, np.inf, np.nan])
@pytest.mark.parametrize("precision,bessel", [(capi.PREC_FP64, capi.BESSEL_AUTO), (capi.PREC_FP64, capi.BESSEL_EXACT),
                                              (capi.PREC_MIXED, capi.BESSEL_EXACT)])
def test_bad_point_refused_before_contraction(handle, bad_r, precision, bessel):
    # one bad radius in a catalogue large enough for the reference table (N > 1000): the call
    # fails with the domain error and reports zero phase times, i.e. no contraction ran
    pts = workloads.uniform_ball(5000, 300.0, seed=41)
    pts[1234, 0] = bad_r
    print(f"case: bad radius {bad_r}, precision {precision}, bessel {bessel}", flush=True)
    d = torch.from_numpy(pts).cuda()
    with pytest.raises(ValueError, match="r < 0 or a non-finite"):
        capi.sfb_direct_torch(handle, d, 24, 12, 300.0, precision=precision, bessel=bessel)
    assert handle.last_phase_ms() == (0.0, 0.0, 0.0)
    # batched: the bad catalogue among good ones
    good = torch.from_numpy(workloads.uniform_ball(3000, 500.0, seed=42)).cuda()
    cats = [dict(points=good, lmax=20, nmax=10, rmax=500.0), dict(points=d, lmax=24, nmax=12, rmax=300.0)]
    with pytest.raises(ValueError, match="r < 0 or a non-finite"):
        capi.sfb_direct_batched_torch(handle, cats, precision=precision)
    assert handle.last_phase_ms() == (0.0, 0.0, 0.0)
    # the handle stays usable
    _, c, _ = capi.sfb_direct_torch(handle, good, 20, 10, 500.0, precision=precision, bessel=bessel)
    assert torch.isfinite(c).all()


def test_healpix_bad_csr_refused(handle):
    # a caller-built CSR whose radii or offsets are invalid is refused before the contraction
    nside = 4
    npix = 12 * nside * nside
    pts = workloads.uniform_ball(2000, 300.0, seed=43)
    off, rad = capi.pixelize_torch(handle, torch.from_numpy(pts).cuda(), nside)
    torch.cuda.synchronize()
    for what, mutate, exc, msg in [
            ("negative radius", lambda o, r: r.__setitem__(7, -5.0), ValueError, "r < 0"),
            ("nan radius", lambda o, r: r.__setitem__(9, float("nan")), ValueError, "r < 0"),
            ("decreasing offsets", lambda o, r: o.__setitem__(5, o[6] + 1), ValueError, "CSR"),
            ("offset 0 not zero", lambda o, r: o.__setitem__(0, 1), ValueError, "CSR")]:
        o, r = off.clone(), rad.clone()
        mutate(o, r)
        print(f"case: healpix {what}", flush=True)
        with pytest.raises(exc, match=msg):
            capi.sfb_healpix_torch(handle, o, r, nside, 16, 8, 300.0)
        assert handle.last_phase_ms() == (0.0, 0.0, 0.0)
    k, c, _ = capi.sfb_healpix_torch(handle, off, rad, nside, 16, 8, 300.0)
    assert torch.isfinite(c).all() and off.shape[0] == npix + 1
